// nebo_stencil.hip -- hand-written kernels for the staggered stencil chains.
//
// Face values are computed exactly as the reference's recursive evaluation does
// (spatialops/NeboStencils.h:407-448: value*coef, left fold in point order), so
// reusing a face value for the two cells that share it gives the same bits as
// recomputing it.  Only the order of memory accesses differs.
//
// Layout: every field is one x-fastest block; the kernels take the frame origin
// (window offset + valid ghost minus) and the shared strides X = glob[0],
// XY = glob[0]*glob[1].
//
// Work split: a workgroup of 64*TY threads (one wave per y row) covers a 64 (x)
// by TY (y) column tile and marches through a chunk of ZC z planes, keeping the
// column's z neighbours and the shared z face in registers.  Two schedules:
//
//  * direct (variant 0): x and y neighbours are plain loads served by L1/L2;
//    every plane costs 10 loads per cell and the wave waits for them before it
//    computes the plane.
//  * staged (variants 1, 2): each thread loads its own cell of the next planes
//    ahead of use (1 or 2 planes of prefetch held in registers), x neighbours
//    come from the adjacent lanes through DPP wave shifts, y neighbours from the
//    plane's rows staged in LDS (double-buffered, one barrier per plane).  Three
//    loads per field and plane: the cell, an x-edge load that the two end lanes
//    use for the tile's x halo, and a row load that waves 0 and TY-1 use for the
//    y halo rows (for the other lanes and waves these two re-read their own cell,
//    so every load is unconditional and the prefetch stays in flight).
//
// Halo rows and planes are re-read by the neighbouring tile, so larger TY and
// ZC cut the traffic beyond L2 ((TY+2)/TY * (ZC+2)/ZC for the operands) at the
// cost of fewer workgroups; the best shape and schedule are measured on the
// device (nebo_stencil_autotune).  Workgroups are numbered so that consecutive
// tiles, which share halo rows, land on the same XCD and therefore the same L2.

#include "nebo_internal.h"
#include "nebo_device.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <sstream>
#include <utility>
#include <vector>

namespace nebo {
namespace {

constexpr int kTx = 64;

// Where one thread of a staged tile loads within a plane.  Every coordinate is
// clamped into [lo-1, hi] on its axis, the box the one-cell stencil reads, so a
// thread outside the limits still loads in bounds (its values are never used).
struct StageGeom {
  int oc;        // own cell (x, y), clamped
  int oe;        // lane 0: (x0-1, y); lane 63: (x0+64, y); other lanes: own cell
  int oh;        // wave 0: (x, y0-1); wave TY-1: (x, y0+TY); other waves: own cell
  int halo_row;  // LDS row the halo load goes to (0 or TY+1), -1 for other waves
};

template <int TY>
__device__ __forceinline__ StageGeom stage_geom(const Tiles& T, int X, int x0, int y0, int lane, int wy)
{
  const int xc = min(x0 + lane, T.hi[0]), yc = min(y0 + wy, T.hi[1]);
  const int xe = lane == 0 ? x0 - 1 : lane == kTx - 1 ? min(x0 + kTx, T.hi[0]) : xc;
  const int yh = wy == 0 ? y0 - 1 : wy == TY - 1 ? min(y0 + TY, T.hi[1]) : yc;
  StageGeom g;
  g.oc = xc + X * yc;
  g.oe = xe + X * yc;
  g.oh = xc + X * yh;
  g.halo_row = wy == 0 ? 0 : wy == TY - 1 ? TY + 1 : -1;
  return g;
}

// ---- div(interp(k) * grad(phi)) ------------------------------------------------

// FORM 0: (Dx + Dy) + Dz;  FORM 1: ((-Dx) - Dy) - Dz;  FORM 2: phi + dt*((Dx + Dy) + Dz)
template <int FORM>
__device__ __forceinline__ double divk_combine(double Dx, double Dy, double Dz, double pc, double dt)
{
  if constexpr (FORM == 0) return (Dx + Dy) + Dz;
  else if constexpr (FORM == 1) return ((-Dx) - Dy) - Dz;
  else return pc + dt * ((Dx + Dy) + Dz);   // explicit Euler step of the diffusion equation
}

struct DivkCoefs {
  double ix0, ix1, gx0, gx1, iy0, iy1, gy0, gy1, iz0, iz1, gz0, gz1;
  double dx0, dx1, dy0, dy1, dz0, dz1;
};

__device__ __forceinline__ DivkCoefs divk_coefs(const nebo_axis_coefs_t& c)
{
  return DivkCoefs{c.interp[0][0], c.interp[0][1], c.grad[0][0], c.grad[0][1],
                   c.interp[1][0], c.interp[1][1], c.grad[1][0], c.grad[1][1],
                   c.interp[2][0], c.interp[2][1], c.grad[2][0], c.grad[2][1],
                   c.div[0][0],    c.div[0][1],    c.div[1][0],  c.div[1][1],
                   c.div[2][0],    c.div[2][1]};
}

template <int FORM, int TY, int ZC>
__global__ __launch_bounds__(kTx * TY) void divkgrad_kernel(
    double* __restrict__ r, const double* __restrict__ k, const double* __restrict__ p,
    Tiles T, int X, int XY, nebo_axis_coefs_t cc, double dt)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int x = T.lo[0] + tx * kTx + (threadIdx.x & (kTx - 1));
  const int y = T.lo[1] + ty * TY + (threadIdx.x / kTx);
  if (x >= T.hi[0] || y >= T.hi[1]) return;
  const int z0 = T.lo[2] + tz * ZC;
  const int z1 = min(z0 + ZC, T.hi[2]);
  const DivkCoefs c = divk_coefs(cc);

  int i = x + X * y + XY * z0;
  double pc = p[i], kc = k[i];
  double fz_lo = face_flux(k[i - XY], kc, p[i - XY], pc, c.iz0, c.iz1, c.gz0, c.gz1);
  for (int z = z0; z < z1; ++z, i += XY) {
    const double pz = p[i + XY], kz = k[i + XY];
    const double pxm = p[i - 1], pxp = p[i + 1], kxm = k[i - 1], kxp = k[i + 1];
    const double pym = p[i - X], pyp = p[i + X], kym = k[i - X], kyp = k[i + X];
    const double fx_lo = face_flux(kxm, kc, pxm, pc, c.ix0, c.ix1, c.gx0, c.gx1);
    const double fx_hi = face_flux(kc, kxp, pc, pxp, c.ix0, c.ix1, c.gx0, c.gx1);
    const double fy_lo = face_flux(kym, kc, pym, pc, c.iy0, c.iy1, c.gy0, c.gy1);
    const double fy_hi = face_flux(kc, kyp, pc, pyp, c.iy0, c.iy1, c.gy0, c.gy1);
    const double fz_hi = face_flux(kc, kz, pc, pz, c.iz0, c.iz1, c.gz0, c.gz1);
    const double Dx = fx_lo * c.dx0 + fx_hi * c.dx1;
    const double Dy = fy_lo * c.dy0 + fy_hi * c.dy1;
    const double Dz = fz_lo * c.dz0 + fz_hi * c.dz1;
    __builtin_nontemporal_store(divk_combine<FORM>(Dx, Dy, Dz, pc, dt), &r[i]);
    pc = pz;
    kc = kz;
    fz_lo = fz_hi;
  }
}

// Staged schedule, PD = planes of prefetch (1 or 2).
template <int FORM, int TY, int ZC, int PD>
__global__ __launch_bounds__(kTx * TY) void divkgrad_staged_kernel(
    double* __restrict__ r, const double* __restrict__ k, const double* __restrict__ p,
    Tiles T, int X, int XY, nebo_axis_coefs_t cc, double dt)
{
  static_assert(TY >= 2 && (PD == 1 || PD == 2), "staged tiles need two rows and a prefetch of 1 or 2");
  __shared__ double sp[2][TY + 2][kTx];
  __shared__ double sk[2][TY + 2][kTx];

  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int lane = threadIdx.x & (kTx - 1);
  const int wy = threadIdx.x / kTx;
  const int x0 = T.lo[0] + tx * kTx, y0 = T.lo[1] + ty * TY;
  const int x = x0 + lane, y = y0 + wy;
  const bool active = x < T.hi[0] && y < T.hi[1];
  const int z0 = T.lo[2] + tz * ZC;
  const int z1 = min(z0 + ZC, T.hi[2]);
  const StageGeom g = stage_geom<TY>(T, X, x0, y0, lane, wy);
  const DivkCoefs c = divk_coefs(cc);

  struct Plane { double p, k, pe, ke, ph, kh; };
  // plane z, any z in [z0-1, z1] (inside the readable box)
  auto load = [&](int z) {
    const int b = XY * z;
    return Plane{p[b + g.oc], k[b + g.oc], p[b + g.oe], k[b + g.oe], p[b + g.oh], k[b + g.oh]};
  };
  auto stage = [&](int buf, const Plane& v) {
    sp[buf][wy + 1][lane] = v.p;
    sk[buf][wy + 1][lane] = v.k;
    if (g.halo_row >= 0) {
      sp[buf][g.halo_row][lane] = v.ph;
      sk[buf][g.halo_row][lane] = v.kh;
    }
  };

  // Register slots: at step z (r = z - z0) slot r % P holds plane z and slot
  // (r + j) % P plane z + j.  One period of P steps is unrolled so the slots are
  // renamed rather than copied: a copy would wait for its load and undo the
  // prefetch.  The chunk is padded to whole periods; padded steps load clamped
  // planes and store nothing.
  constexpr int P = PD + 1;
  Plane S[P];
  S[0] = load(z0);
  const double pzm = p[XY * (z0 - 1) + g.oc], kzm = k[XY * (z0 - 1) + g.oc];
#pragma unroll
  for (int j = 1; j < P; ++j) S[j] = load(min(z0 + j, z1));
  stage(0, S[0]);
  __syncthreads();

  double fz_lo = face_flux(kzm, S[0].k, pzm, S[0].p, c.iz0, c.iz1, c.gz0, c.gz1);
  int buf = 0;
  const int periods = (z1 - z0 + P - 1) / P;
  for (int q = 0; q < periods; ++q) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int z = z0 + q * P + j;
      const double pc = S[j].p, kc = S[j].k, pe = S[j].pe, ke = S[j].ke;
      const Plane& B1 = S[(j + 1) % P];   // plane z+1
      const double pl = from_lower_lane(pc), pu = from_upper_lane(pc);
      const double kl = from_lower_lane(kc), ku = from_upper_lane(kc);
      const double pxm = lane == 0 ? pe : pl, pxp = lane == kTx - 1 ? pe : pu;
      const double kxm = lane == 0 ? ke : kl, kxp = lane == kTx - 1 ? ke : ku;
      const double pym = sp[buf][wy][lane], pyp = sp[buf][wy + 2][lane];
      const double kym = sk[buf][wy][lane], kyp = sk[buf][wy + 2][lane];
      const double fx_lo = face_flux(kxm, kc, pxm, pc, c.ix0, c.ix1, c.gx0, c.gx1);
      const double fx_hi = face_flux(kc, kxp, pc, pxp, c.ix0, c.ix1, c.gx0, c.gx1);
      const double fy_lo = face_flux(kym, kc, pym, pc, c.iy0, c.iy1, c.gy0, c.gy1);
      const double fy_hi = face_flux(kc, kyp, pc, pyp, c.iy0, c.iy1, c.gy0, c.gy1);
      const double fz_hi = face_flux(kc, B1.k, pc, B1.p, c.iz0, c.iz1, c.gz0, c.gz1);
      const double Dx = fx_lo * c.dx0 + fx_hi * c.dx1;
      const double Dy = fy_lo * c.dy0 + fy_hi * c.dy1;
      const double Dz = fz_lo * c.dz0 + fz_hi * c.dz1;
      const double v = divk_combine<FORM>(Dx, Dy, Dz, pc, dt);
      if (active && z < z1) __builtin_nontemporal_store(v, &r[x + X * y + XY * z]);
      // plane z is consumed: its slot receives plane z+P (same registers, no copy)
      S[j] = load(min(z + P, z1));
      stage(buf ^ 1, B1);   // plane z+1 for the next step
      __syncthreads();
      buf ^= 1;
      fz_lo = fz_hi;
    }
  }
}

// Column schedule (variant 3): W waves side by side in x (a 64*W-wide strip, so
// the loads in flight cover whole contiguous rows), each lane one x column of RY
// rows, the wave marching ZC planes.  No LDS and no barrier: y neighbours are the
// lane's own rows plus one halo row on each side, x neighbours come from the
// adjacent lanes (DPP) and at the wave's ends from the edge load (readlane).
// Three register slots hold planes z, z+1 and z+2; once plane z is done its slot
// receives plane z+3, so loads are issued two planes ahead.  One period of three
// steps is unrolled (slots are renamed, never copied); the one or two steps left
// after the whole periods run once after the loop, so a chunk of any length does
// no padded work.  The chunk length sets the number of waves (for a 256^3 box at
// 4 rows per lane, a 16-plane chunk gives 4096 waves, the whole GPU's residency
// at 4 waves per SIMD), which hides the memory latency.
//
// XR (x-face reuse): each lane computes only its high x face; its low face is the
// lower lane's high face (DPP).  The faces at the wave's two ends are computed for
// all RY rows at once, once per plane: lane j (< RY) takes the left face of row j
// from its edge load and an inner load (x = xw0), lane 32+j the right face from
// its edge load and an inner load (x = xw0+63); each row reads its two end faces
// back with readlane.  A face has the same value wherever it is computed.
//
// NW waves stacked in y per workgroup: a wave's halo rows are its y neighbours'
// own rows, loaded at the same time on the same CU, so they hit L1; a halo row
// that only another CU loaded comes back through L2 at about the cost of an HBM
// row.  (Measured at 256^3 with 4 rows per lane: the halo rows cost ~13 us of an
// ~84 us launch with one wave per workgroup, ~3 us with four stacked.)

template <int FORM, int W, int NW, int RY, int ZC, bool XR, int PF = 2, int SP = 2>
__global__ __launch_bounds__(kTx * W * NW) void divkgrad_column_kernel(
    double* __restrict__ r, const double* __restrict__ k, const double* __restrict__ p,
    Tiles T, int X, int XY, nebo_axis_coefs_t cc, double dt)
{
  static_assert(RY >= 1 && RY <= 16, "rows per lane");
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int lane = threadIdx.x & (kTx - 1);
  const int wave = threadIdx.x / kTx;
  const int xw0 = T.lo[0] + tx * (kTx * W) + (wave % W) * kTx;
  const int x = xw0 + lane;
  const int y0 = T.lo[1] + ty * (RY * NW) + (wave / W) * RY;
  const int z0 = T.lo[2] + tz * ZC;
  const int z1 = min(z0 + ZC, T.hi[2]);
  const bool xin = x < T.hi[0];
  const ColumnGeom<RY> g = column_geom<RY>(T, X, xw0, y0, lane);
  const DivkCoefs c = divk_coefs(cc);

  struct Col { double p[RY + 2], k[RY + 2], pe, ke, pi, ki; };
  auto load = [&](int z) {
    Col v;
    const Rsrc pb = plane_rsrc(plane_base(p, T, X, XY, z), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z), X, XY);
#pragma unroll
    for (int q = 0; q < RY + 2; ++q) {
      v.p[q] = load_at(pb, g.row[q]);
      v.k[q] = load_at(kb, g.row[q]);
    }
    v.pe = load_at(pb, g.edge);
    v.ke = load_at(kb, g.edge);
    if constexpr (XR) {
      v.pi = load_at(pb, g.inner);
      v.ki = load_at(kb, g.inner);
    }
    return v;
  };

  constexpr int NS = PF + 1;   // register slots: planes z .. z+PF
  Col S[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) S[q] = load(min(z0 + q, z1));
  double fz[RY];
  {
    const Rsrc pb = plane_rsrc(plane_base(p, T, X, XY, z0 - 1), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z0 - 1), X, XY);
#pragma unroll
    for (int j = 0; j < RY; ++j)
      fz[j] = face_flux(load_at(kb, g.row[j + 1]), S[0].k[j + 1], load_at(pb, g.row[j + 1]), S[0].p[j + 1], c.iz0,
                   c.iz1, c.gz0, c.gz1);
  }
  // one step: plane z, whose slot is jj (a compile-time index)
  auto step = [&](auto JJ, const int z) {
    constexpr int jj = decltype(JJ)::value;
    {
      const Col& C0 = S[jj];              // plane z
      const Col& C1 = S[(jj + 1) % NS];   // plane z+1
      // the y face between rows j-1 and j is computed once, as row j-1's high face
      double fy_lo = face_flux(C0.k[0], C0.k[1], C0.p[0], C0.p[1], c.iy0, c.iy1, c.gy0, c.gy1);
      // XR: the wave's end faces of every row (lane j: left face of row j, lane 32+j: right face)
      double fend = 0.0;
      if constexpr (XR) {
        const bool left = lane < 32;
        fend = face_flux(left ? C0.ke : C0.ki, left ? C0.ki : C0.ke, left ? C0.pe : C0.pi, left ? C0.pi : C0.pe,
                    c.ix0, c.ix1, c.gx0, c.gx1);
      }
#pragma unroll
      for (int j = 0; j < RY; ++j) {
        const double pc = C0.p[j + 1], kc = C0.k[j + 1];
        double fx_lo, fx_hi;
        if constexpr (XR) {
          const double pu = from_upper_lane(pc), ku = from_upper_lane(kc);
          const double own = face_flux(kc, ku, pc, pu, c.ix0, c.ix1, c.gx0, c.gx1);
          // lane reads outside the selects: every lane takes them (see lane_value)
          const double end_hi = lane_value(fend, 32 + j), end_lo = lane_value(fend, j);
          fx_hi = lane == kTx - 1 ? end_hi : own;
          const double below = from_lower_lane(fx_hi);
          fx_lo = lane == 0 ? end_lo : below;
        } else {
          const double pl = from_lower_lane(pc), pu = from_upper_lane(pc);
          const double kl = from_lower_lane(kc), ku = from_upper_lane(kc);
          const double pel = lane_value(C0.pe, j), per = lane_value(C0.pe, 32 + j);
          const double kel = lane_value(C0.ke, j), ker = lane_value(C0.ke, 32 + j);
          const double pxm = lane == 0 ? pel : pl, pxp = lane == kTx - 1 ? per : pu;
          const double kxm = lane == 0 ? kel : kl, kxp = lane == kTx - 1 ? ker : ku;
          fx_lo = face_flux(kxm, kc, pxm, pc, c.ix0, c.ix1, c.gx0, c.gx1);
          fx_hi = face_flux(kc, kxp, pc, pxp, c.ix0, c.ix1, c.gx0, c.gx1);
        }
        const double fy_hi = face_flux(kc, C0.k[j + 2], pc, C0.p[j + 2], c.iy0, c.iy1, c.gy0, c.gy1);
        const double fz_hi = face_flux(kc, C1.k[j + 1], pc, C1.p[j + 1], c.iz0, c.iz1, c.gz0, c.gz1);
        const double Dx = fx_lo * c.dx0 + fx_hi * c.dx1;
        const double Dy = fy_lo * c.dy0 + fy_hi * c.dy1;
        const double Dz = fz[j] * c.dz0 + fz_hi * c.dz1;
        fz[j] = fz_hi;
        fy_lo = fy_hi;
        const double v = divk_combine<FORM>(Dx, Dy, Dz, pc, dt);
        if (xin && y0 + j < T.hi[1]) store_at<SP>(plane_rsrc(plane_base(r, T, X, XY, z), X, XY), g.row[j + 1], v);
      }
      S[jj] = load(min(z + NS, z1));      // plane z is consumed: its slot takes z+NS
    }
  };
  // whole periods of NS steps, then the steps left over
  const int full = (z1 - z0) / NS, rem = (z1 - z0) - NS * full;
  int z = z0;
  for (int q = 0; q < full; ++q, z += NS)
    unroll_steps([&](auto J) { step(J, z + decltype(J)::value); }, std::make_integer_sequence<int, NS>{});
  unroll_steps([&](auto J) { if (rem > decltype(J)::value) step(J, z + decltype(J)::value); },
               std::make_integer_sequence<int, NS - 1>{});
}

// Pair-column schedule (variants 10-16): the column schedule with two adjacent
// columns per lane, so every load and store moves 16 bytes and a wave covers
// 128 columns (W waves side by side, NW stacked in y, RY rows per lane, loads PF
// planes ahead).  A lane holding columns x, x+1 takes x-1 from the lower lane's
// second column and x+2 from the upper lane's first (DPP); the wave's two outer
// neighbours come from one edge load per field (readlane).  Each lane forms the
// three x faces of its two cells; faces and cells use the reference's
// operations in its order, so results equal the other schedules bit for bit.
// Used only on boxes whose width is a multiple of 128 W (usable_shape).
template <int FORM, int W, int NW, int RY, int ZC, int PF>
__global__ __launch_bounds__(kTx * W * NW) void divkgrad_pair_kernel(
    double* __restrict__ r, const double* __restrict__ k, const double* __restrict__ p,
    Tiles T, int X, int XY, nebo_axis_coefs_t cc, double dt)
{
  static_assert(RY >= 1 && RY <= 8, "rows per lane");
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int lane = threadIdx.x & (kTx - 1);
  const int wave = threadIdx.x / kTx;
  const int xw0 = T.lo[0] + tx * (2 * kTx * W) + (wave % W) * (2 * kTx);
  const int y0 = T.lo[1] + ty * (RY * NW) + (wave / W) * RY;
  const int z0 = T.lo[2] + tz * ZC;
  const int z1 = min(z0 + ZC, T.hi[2]);
  const PairGeom<RY> g = pair_geom<RY>(T, X, xw0, y0, lane);
  const DivkCoefs c = divk_coefs(cc);

  struct Col { D2 p[RY + 2], k[RY + 2]; double pe, ke; };
  auto load = [&](int z) {
    Col v;
    const Rsrc pb = plane_rsrc(plane_base(p, T, X, XY, z), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z), X, XY);
#pragma unroll
    for (int q = 0; q < RY + 2; ++q) {
      v.p[q] = load2_at(pb, g.row[q]);
      v.k[q] = load2_at(kb, g.row[q]);
    }
    v.pe = load_at(pb, g.edge);
    v.ke = load_at(kb, g.edge);
    return v;
  };

  constexpr int NS = PF + 1;   // register slots: planes z .. z+PF
  Col S[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) S[q] = load(min(z0 + q, z1));
  double fz[RY][2];
  {
    const Rsrc pb = plane_rsrc(plane_base(p, T, X, XY, z0 - 1), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z0 - 1), X, XY);
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const D2 pm = load2_at(pb, g.row[j + 1]), km = load2_at(kb, g.row[j + 1]);
#pragma unroll
      for (int h = 0; h < 2; ++h)
        fz[j][h] = face_flux(km[h], S[0].k[j + 1][h], pm[h], S[0].p[j + 1][h], c.iz0, c.iz1, c.gz0, c.gz1);
    }
  }
  auto step = [&](auto JJ, const int z) {
    constexpr int jj = decltype(JJ)::value;
    const Col& C0 = S[jj];              // plane z
    const Col& C1 = S[(jj + 1) % NS];   // plane z+1
    double fy_lo[2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
      fy_lo[h] = face_flux(C0.k[0][h], C0.k[1][h], C0.p[0][h], C0.p[1][h], c.iy0, c.iy1, c.gy0, c.gy1);
    const Rsrc rb = plane_rsrc(plane_base(r, T, X, XY, z), X, XY);
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const D2 a = C0.p[j + 1], b = C0.k[j + 1];
      // x neighbours: the exchanges are taken by every lane, then the edge lanes
      // replace theirs with the edge values
      const double pl_d = from_lower_lane(a[1]), kl_d = from_lower_lane(b[1]);
      const double pr_d = from_upper_lane(a[0]), kr_d = from_upper_lane(b[0]);
      const double pel = lane_value(C0.pe, j), kel = lane_value(C0.ke, j);
      const double per = lane_value(C0.pe, 32 + j), ker = lane_value(C0.ke, 32 + j);
      const double pl = lane == 0 ? pel : pl_d, kl = lane == 0 ? kel : kl_d;
      const double pr = lane == kTx - 1 ? per : pr_d, kr = lane == kTx - 1 ? ker : kr_d;
      const double F0 = face_flux(kl, b[0], pl, a[0], c.ix0, c.ix1, c.gx0, c.gx1);
      const double F1 = face_flux(b[0], b[1], a[0], a[1], c.ix0, c.ix1, c.gx0, c.gx1);
      const double F2 = face_flux(b[1], kr, a[1], pr, c.ix0, c.ix1, c.gx0, c.gx1);
      const double Dx[2] = {F0 * c.dx0 + F1 * c.dx1, F1 * c.dx0 + F2 * c.dx1};
      D2 v;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double fy_hi = face_flux(b[h], C0.k[j + 2][h], a[h], C0.p[j + 2][h], c.iy0, c.iy1, c.gy0, c.gy1);
        const double fz_hi = face_flux(b[h], C1.k[j + 1][h], a[h], C1.p[j + 1][h], c.iz0, c.iz1, c.gz0, c.gz1);
        const double Dy = fy_lo[h] * c.dy0 + fy_hi * c.dy1;
        const double Dz = fz[j][h] * c.dz0 + fz_hi * c.dz1;
        fz[j][h] = fz_hi;
        fy_lo[h] = fy_hi;
        v[h] = divk_combine<FORM>(Dx[h], Dy, Dz, a[h], dt);
      }
      if (y0 + j < T.hi[1]) store2_nt_at(rb, g.row[j + 1], v);
    }
    S[jj] = load(min(z + NS, z1));      // plane z is consumed: its slot takes z+NS
  };
  // whole periods of NS steps, then the steps left over
  const int full = (z1 - z0) / NS, rem = (z1 - z0) - NS * full;
  int z = z0;
  for (int q = 0; q < full; ++q, z += NS)
    unroll_steps([&](auto J) { step(J, z + decltype(J)::value); }, std::make_integer_sequence<int, NS>{});
  unroll_steps([&](auto J) { if (rem > decltype(J)::value) step(J, z + decltype(J)::value); },
               std::make_integer_sequence<int, NS - 1>{});
}

// ---- phi + s*lapl(phi) -----------------------------------------------------------

// out = phi + s*((Lx + Ly) + Lz),  L_d = (phi[-1]*g0 + phi[0]*g1)*d0 + (phi[0]*g0 + phi[+1]*g1)*d1
struct LaplCoefs {
  double gx0, gx1, gy0, gy1, gz0, gz1, dx0, dx1, dy0, dy1, dz0, dz1;
};

__device__ __forceinline__ LaplCoefs lapl_coefs(const nebo_axis_coefs_t& c)
{
  return LaplCoefs{c.grad[0][0], c.grad[0][1], c.grad[1][0], c.grad[1][1], c.grad[2][0], c.grad[2][1],
                   c.div[0][0],  c.div[0][1],  c.div[1][0],  c.div[1][1],  c.div[2][0],  c.div[2][1]};
}

__device__ __forceinline__ double lapl_cell(double pc, double pxm, double pxp, double pym, double pyp,
                                            double gz_lo, double gz_hi, const LaplCoefs& c, double s)
{
  const double Lx = (pxm * c.gx0 + pc * c.gx1) * c.dx0 + (pc * c.gx0 + pxp * c.gx1) * c.dx1;
  const double Ly = (pym * c.gy0 + pc * c.gy1) * c.dy0 + (pc * c.gy0 + pyp * c.gy1) * c.dy1;
  const double Lz = gz_lo * c.dz0 + gz_hi * c.dz1;
  return pc + s * ((Lx + Ly) + Lz);
}

template <int TY, int ZC>
__global__ __launch_bounds__(kTx * TY) void lapl_euler_kernel(
    double* __restrict__ out, const double* __restrict__ p, Tiles T, int X, int XY,
    nebo_axis_coefs_t cc, double s)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int x = T.lo[0] + tx * kTx + (threadIdx.x & (kTx - 1));
  const int y = T.lo[1] + ty * TY + (threadIdx.x / kTx);
  if (x >= T.hi[0] || y >= T.hi[1]) return;
  const int z0 = T.lo[2] + tz * ZC;
  const int z1 = min(z0 + ZC, T.hi[2]);
  const LaplCoefs c = lapl_coefs(cc);

  int i = x + X * y + XY * z0;
  double pc = p[i];
  double gz_lo = p[i - XY] * c.gz0 + pc * c.gz1;
  for (int z = z0; z < z1; ++z, i += XY) {
    const double pz = p[i + XY];
    const double pxm = p[i - 1], pxp = p[i + 1], pym = p[i - X], pyp = p[i + X];
    const double gz_hi = pc * c.gz0 + pz * c.gz1;
    __builtin_nontemporal_store(lapl_cell(pc, pxm, pxp, pym, pyp, gz_lo, gz_hi, c, s), &out[i]);
    pc = pz;
    gz_lo = gz_hi;
  }
}

template <int TY, int ZC, int PD>
__global__ __launch_bounds__(kTx * TY) void lapl_euler_staged_kernel(
    double* __restrict__ out, const double* __restrict__ p, Tiles T, int X, int XY,
    nebo_axis_coefs_t cc, double s)
{
  static_assert(TY >= 2 && (PD == 1 || PD == 2), "staged tiles need two rows and a prefetch of 1 or 2");
  __shared__ double sp[2][TY + 2][kTx];

  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int lane = threadIdx.x & (kTx - 1);
  const int wy = threadIdx.x / kTx;
  const int x0 = T.lo[0] + tx * kTx, y0 = T.lo[1] + ty * TY;
  const int x = x0 + lane, y = y0 + wy;
  const bool active = x < T.hi[0] && y < T.hi[1];
  const int z0 = T.lo[2] + tz * ZC;
  const int z1 = min(z0 + ZC, T.hi[2]);
  const StageGeom g = stage_geom<TY>(T, X, x0, y0, lane, wy);
  const LaplCoefs c = lapl_coefs(cc);

  struct Plane { double p, pe, ph; };
  auto load = [&](int z) {
    const int b = XY * z;
    return Plane{p[b + g.oc], p[b + g.oe], p[b + g.oh]};
  };
  auto stage = [&](int buf, const Plane& v) {
    sp[buf][wy + 1][lane] = v.p;
    if (g.halo_row >= 0) sp[buf][g.halo_row][lane] = v.ph;
  };

  // register slots renamed by an unrolled period, as in divkgrad_staged_kernel
  constexpr int P = PD + 1;
  Plane S[P];
  S[0] = load(z0);
  const double pzm = p[XY * (z0 - 1) + g.oc];
#pragma unroll
  for (int j = 1; j < P; ++j) S[j] = load(min(z0 + j, z1));
  stage(0, S[0]);
  __syncthreads();

  double gz_lo = pzm * c.gz0 + S[0].p * c.gz1;
  int buf = 0;
  const int periods = (z1 - z0 + P - 1) / P;
  for (int q = 0; q < periods; ++q) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int z = z0 + q * P + j;
      const double pc = S[j].p, pe = S[j].pe;
      const Plane& B1 = S[(j + 1) % P];
      const double pl = from_lower_lane(pc), pu = from_upper_lane(pc);
      const double pxm = lane == 0 ? pe : pl, pxp = lane == kTx - 1 ? pe : pu;
      const double pym = sp[buf][wy][lane], pyp = sp[buf][wy + 2][lane];
      const double gz_hi = pc * c.gz0 + B1.p * c.gz1;
      const double v = lapl_cell(pc, pxm, pxp, pym, pyp, gz_lo, gz_hi, c, s);
      if (active && z < z1) __builtin_nontemporal_store(v, &out[x + X * y + XY * z]);
      S[j] = load(min(z + P, z1));
      stage(buf ^ 1, B1);
      __syncthreads();
      buf ^= 1;
      gz_lo = gz_hi;
    }
  }
}

// Column schedule of the Laplacian + Euler update (see divkgrad_column_kernel).
template <int W, int NW, int RY, int ZC, int SP = 2>
__global__ __launch_bounds__(kTx * W * NW) void lapl_euler_column_kernel(
    double* __restrict__ out, const double* __restrict__ p, Tiles T, int X, int XY,
    nebo_axis_coefs_t cc, double s)
{
  static_assert(RY >= 1 && RY <= 16, "rows per lane");
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int lane = threadIdx.x & (kTx - 1);
  const int wave = threadIdx.x / kTx;
  const int xw0 = T.lo[0] + tx * (kTx * W) + (wave % W) * kTx;
  const int x = xw0 + lane;
  const int y0 = T.lo[1] + ty * (RY * NW) + (wave / W) * RY;
  const int z0 = T.lo[2] + tz * ZC;
  const int z1 = min(z0 + ZC, T.hi[2]);
  const bool xin = x < T.hi[0];
  const ColumnGeom<RY> g = column_geom<RY>(T, X, xw0, y0, lane);
  const LaplCoefs c = lapl_coefs(cc);

  struct Col { double p[RY + 2], pe; };
  auto load = [&](int z) {
    Col v;
    const Rsrc pb = plane_rsrc(plane_base(p, T, X, XY, z), X, XY);
#pragma unroll
    for (int q = 0; q < RY + 2; ++q) v.p[q] = load_at(pb, g.row[q]);
    v.pe = load_at(pb, g.edge);
    return v;
  };

  Col S[3];
  S[0] = load(z0);
  S[1] = load(min(z0 + 1, z1));
  S[2] = load(min(z0 + 2, z1));
  double gz[RY];
#pragma unroll
  for (int j = 0; j < RY; ++j)
    gz[j] = load_at(plane_rsrc(plane_base(p, T, X, XY, z0 - 1), X, XY), g.row[j + 1]) * c.gz0 + S[0].p[j + 1] * c.gz1;
  auto step = [&](auto JJ, const int z) {
    constexpr int jj = decltype(JJ)::value;
    const Col& C0 = S[jj];
    const Col& C1 = S[(jj + 1) % 3];
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const double pc = C0.p[j + 1];
      const double pl = from_lower_lane(pc), pu = from_upper_lane(pc);
      const double pel = lane_value(C0.pe, j), per = lane_value(C0.pe, 32 + j);
      const double pxm = lane == 0 ? pel : pl, pxp = lane == kTx - 1 ? per : pu;
      const double gz_hi = pc * c.gz0 + C1.p[j + 1] * c.gz1;
      const double v = lapl_cell(pc, pxm, pxp, C0.p[j], C0.p[j + 2], gz[j], gz_hi, c, s);
      gz[j] = gz_hi;
      if (xin && y0 + j < T.hi[1]) store_at<SP>(plane_rsrc(plane_base(out, T, X, XY, z), X, XY), g.row[j + 1], v);
    }
    S[jj] = load(min(z + 3, z1));
  };
  const int full = (z1 - z0) / 3, rem = (z1 - z0) - 3 * full;
  int z = z0;
  for (int q = 0; q < full; ++q, z += 3) {
    step(std::integral_constant<int, 0>{}, z);
    step(std::integral_constant<int, 1>{}, z + 1);
    step(std::integral_constant<int, 2>{}, z + 2);
  }
  if (rem > 0) step(std::integral_constant<int, 0>{}, z);
  if (rem > 1) step(std::integral_constant<int, 1>{}, z + 1);
}

// ---- tile table -------------------------------------------------------------

// variant 0: direct loads; 1, 2: LDS-staged with 1 or 2 planes of prefetch;
// 3: column schedule (ty = rows per lane, tx = 64 * waves side by side);
// 4: column schedule with x-face reuse (the Laplacian kernel has no x-face
//    reuse and runs its column schedule for both 3 and 4);
// 5, 6: column schedule with x-face reuse, 4 resp. 2 rows per lane, ty/4 resp.
//    ty/2 waves stacked in y per workgroup (ty = the workgroup's rows);
// 7, 8: as 5 with loads issued 3 resp. 4 planes ahead instead of 2;
// 9: as 7 with write-back stores instead of nontemporal ones;
// 10-16: pair columns (two adjacent columns per lane, 16-byte accesses,
//    tx = 128 x waves side by side), ty/rows waves stacked in y, with rows per
//    lane and planes of load prefetch (kPairRows, kPairPrefetch):
//      10: 4 rows, 3 planes   11: 2, 3   12: 4, 2   13: 2, 4   14: 2, 2
//      15: 1, 3               16: 1, 4
//    (div(k grad phi); the Laplacian runs its column schedule over the same
//    tile with the same rows per lane).
struct TileShape { int tx, ty, zc, variant; };
constexpr TileShape kShapes[] = {
    {64, 4, 16, 0},   {64, 8, 16, 0},
    {64, 8, 16, 1},   {64, 8, 64, 1},   {64, 16, 32, 1},
    {64, 8, 24, 2},
    {64, 4, 16, 3},   {64, 4, 24, 3},   {64, 2, 32, 3},   {64, 2, 8, 3},    {64, 4, 8, 3},    {64, 2, 4, 3},
    {64, 1, 8, 3},    {64, 1, 16, 3},   {64, 2, 16, 3},
    {64, 4, 16, 4},   {64, 4, 24, 4},   {64, 4, 32, 4},   {64, 2, 32, 4},   {256, 4, 16, 4},
    {256, 4, 64, 4},
    {64, 8, 32, 5},   {64, 16, 16, 5},  {64, 16, 32, 5},  {128, 8, 32, 5},  {128, 8, 64, 5},
    {128, 16, 32, 5},
    {64, 8, 32, 6},   {64, 16, 32, 6},
    {128, 8, 32, 7},  {128, 8, 64, 7},  {64, 16, 64, 7},  {256, 4, 64, 7},  {128, 4, 128, 7},
    {128, 8, 128, 7}, {128, 8, 256, 7}, {128, 8, 512, 7},
    {128, 8, 64, 8},
    {128, 16, 32, 7}, {128, 16, 64, 7}, {256, 8, 64, 7},
    {128, 8, 64, 9},  {128, 8, 512, 9},
    {128, 8, 64, 10}, {128, 8, 32, 10}, {256, 8, 32, 10}, {128, 16, 32, 10}, {128, 8, 512, 10},
    {128, 4, 64, 11}, {128, 8, 64, 11}, {128, 8, 64, 12},
    {128, 8, 32, 11}, {256, 4, 64, 11}, {256, 8, 32, 11}, {128, 16, 32, 11}, {128, 8, 512, 11},
    {128, 8, 64, 13}, {256, 4, 64, 13}, {128, 8, 64, 14}, {128, 8, 64, 15}, {128, 4, 64, 15},
    {128, 8, 64, 16}, {256, 4, 64, 14}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);
// The default tile (rule 4 of select_shape): 128 x 8 rows (2 x 2 waves)
// marching 64 planes with loads 3 planes ahead, the fastest at 256^3 on MI355X
// (one workgroup per CU there).  Planes too large for the column schedule's
// 32-bit offsets run the direct schedule (kFallbackShape).
constexpr int find_shape_c(int tx, int ty, int zc, int variant, int s = 0)
{
  return s >= kNumShapes ? -1
         : (kShapes[s].tx == tx && kShapes[s].ty == ty && kShapes[s].zc == zc && kShapes[s].variant == variant)
             ? s
             : find_shape_c(tx, ty, zc, variant, s + 1);
}
// Pair-column tiles: two columns per lane, 128 columns per wave.
constexpr bool pair_columns(const TileShape& sh) { return sh.variant >= 10; }
constexpr int kPairRows[] = {4, 2, 4, 2, 2, 1, 1};       // variants 10 ..
constexpr int kPairPrefetch[] = {3, 3, 2, 4, 2, 3, 4};
constexpr int kDefaultShape = find_shape_c(128, 8, 64, 7);
// div(k grad phi) on boxes whose width is a multiple of 128: the same tile as
// pair columns, 2 rows per lane (4 waves stacked), loads 3 planes ahead (256^3 on
// MI355X: 66-69 us against 70-73 us for the default tile).
constexpr int kDefaultPairShape = find_shape_c(128, 8, 64, 11);
constexpr int kFallbackShape = 0;
static_assert(kDefaultShape >= 0 && kDefaultPairShape >= 0 && kShapes[kFallbackShape].variant == 0 &&
                  !pair_columns(kShapes[kDefaultShape]),
              "default tiles");

// Rows per lane and waves stacked in y of a column tile.
constexpr int rows_per_lane(const TileShape& sh)
{
  return pair_columns(sh) ? kPairRows[sh.variant - 10]
         : (sh.variant == 5 || sh.variant == 7 || sh.variant == 8 || sh.variant == 9) ? 4
         : sh.variant == 6                                                              ? 2
                                                                                        : sh.ty;
}
// Planes of load prefetch of a column tile.
constexpr int prefetch_planes(const TileShape& sh)
{
  return pair_columns(sh)                        ? kPairPrefetch[sh.variant - 10]
         : (sh.variant == 7 || sh.variant == 9) ? 3
         : sh.variant == 8                       ? 4
                                                 : 2;
}
// Cache policy of a column tile's stores (2 nontemporal, 0 write-back).
constexpr int store_policy(const TileShape& sh) { return sh.variant == 9 ? 0 : 2; }
constexpr int stacked_waves(const TileShape& sh) { return sh.ty / rows_per_lane(sh); }

// Launch geometry of one tile: rows per tile in y and threads per workgroup
// (kernel 0 div(k grad phi), 1 Laplacian + Euler; the Laplacian runs pair tiles
// on its one-column schedule, tx / 64 waves side by side).
struct Geometry { int tile_y, threads; };
constexpr Geometry tile_geometry(const TileShape& sh, int kernel = 0)
{
  return (pair_columns(sh) && kernel == 0) ? Geometry{sh.ty, sh.tx / 2 * stacked_waves(sh)}
         : sh.variant >= 3                 ? Geometry{sh.ty, sh.tx * stacked_waves(sh)}
                                           : Geometry{sh.ty, sh.tx * sh.ty};
}

using DivkFn = void (*)(double*, const double*, const double*, Tiles, int, int, nebo_axis_coefs_t, double);
using LaplFn = void (*)(double*, const double*, Tiles, int, int, nebo_axis_coefs_t, double);

template <int FORM, int S>
constexpr DivkFn divk_fn()
{
  constexpr TileShape sh = kShapes[S];
  static_assert(sh.variant >= 3 || sh.tx == kTx, "only the column schedule spans several waves in x");
  static_assert(sh.ty % rows_per_lane(sh) == 0 && tile_geometry(sh).threads <= 1024, "whole waves, <= 1024 threads");
  static_assert(!pair_columns(sh) || sh.tx % (2 * kTx) == 0, "pair tiles hold whole 128-column waves");
  if constexpr (sh.variant == 0) return divkgrad_kernel<FORM, sh.ty, sh.zc>;
  else if constexpr (pair_columns(sh))
    return divkgrad_pair_kernel<FORM, sh.tx / (2 * kTx), stacked_waves(sh), rows_per_lane(sh), sh.zc,
                                prefetch_planes(sh)>;
  else if constexpr (sh.variant >= 3)
    return divkgrad_column_kernel<FORM, sh.tx / kTx, stacked_waves(sh), rows_per_lane(sh), sh.zc, sh.variant != 3,
                                  prefetch_planes(sh), store_policy(sh)>;
  else return divkgrad_staged_kernel<FORM, sh.ty, sh.zc, sh.variant>;
}

template <int S>
constexpr LaplFn lapl_fn()
{
  constexpr TileShape sh = kShapes[S];
  if constexpr (sh.variant == 0) return lapl_euler_kernel<sh.ty, sh.zc>;
  else if constexpr (sh.variant >= 3)
    return lapl_euler_column_kernel<sh.tx / kTx, stacked_waves(sh), rows_per_lane(sh), sh.zc, store_policy(sh)>;
  else return lapl_euler_staged_kernel<sh.ty, sh.zc, sh.variant>;
}

template <int FORM, size_t... S>
DivkFn divk_pick(int s, std::index_sequence<S...>)
{
  static const DivkFn t[] = {divk_fn<FORM, (int)S>()...};
  return t[s];
}

template <size_t... S>
LaplFn lapl_pick(int s, std::index_sequence<S...>)
{
  static const LaplFn t[] = {lapl_fn<(int)S>()...};
  return t[s];
}

DivkFn divk_table(int form, int s)
{
  using Seq = std::make_index_sequence<kNumShapes>;
  return form == 0 ? divk_pick<0>(s, Seq{}) : form == 1 ? divk_pick<1>(s, Seq{}) : divk_pick<2>(s, Seq{});
}

LaplFn lapl_table(int s) { return lapl_pick(s, std::make_index_sequence<kNumShapes>{}); }

// Tile selection, per kernel (0 div(k grad phi), 1 Laplacian + Euler):
//  1. a tile pinned with nebo_stencil_set_tile applies to every box;
//  2. else the tile nebo_stencil_autotune found fastest for a box of the same
//     dimensions on the same device;
//  3. else, on boxes large enough, the default tile's family (128 x 8, loads 3
//     planes ahead) with the longest z chunk (512, 256 or 128 planes) that still
//     gives every CU a workgroup: one workgroup per CU, all walking the same
//     planes (512^3 on MI355X: 658 us against 775 us with 64-plane chunks, whose
//     eight concurrent z slabs lie 135 MB apart);
//  4. else the default tile when it gives at least half the CUs a workgroup
//     (256^3: 64-plane chunks, one workgroup per CU), for div(k grad phi) on
//     boxes whose width is a multiple of 128 as pair columns;
//  5. smaller boxes run one-wave column tiles of 64 columns x 2 rows per lane
//     with the longest z chunk (16, 8 or 4 planes) that gives every CU two
//     waves (else 4 planes): time steps that alternate two fields, timed in a
//     hipGraph on MI355X, 64^3 3.9 us (Laplacian) and 4.6 us (div(k grad phi))
//     against 7.8 and 9.0 us for the direct schedule's 64 x 4 x 16 tiles;
//     128^3 8.7 and 13.9 us against 10.0 and 18.6 us.
std::atomic<int> g_pinned[2] = {{-1}, {-1}};
std::atomic<int> g_last[2] = {{kDefaultShape}, {kDefaultShape}};   // last pinned or tuned tile
struct TunedTile {
  int kernel, device, n[3], shape;
};
std::mutex g_tuned_mu;
std::vector<TunedTile> g_tuned;

constexpr int kSmallColumn[] = {find_shape_c(64, 2, 16, 3), find_shape_c(64, 2, 8, 3), find_shape_c(64, 2, 4, 3)};
constexpr int kLongChunks[] = {find_shape_c(128, 8, 512, 7), find_shape_c(128, 8, 256, 7),
                               find_shape_c(128, 8, 128, 7)};
static_assert(kSmallColumn[0] >= 0 && kSmallColumn[1] >= 0 && kSmallColumn[2] >= 0 && kLongChunks[0] >= 0 &&
                  kLongChunks[1] >= 0 && kLongChunks[2] >= 0,
              "box-size tiles");


long tile_groups(int shape, const nebo_limits_t& lim)
{
  const TileShape sh = kShapes[shape];
  const Tiles T = make_tiles(lim, sh.tx, tile_geometry(sh).tile_y, sh.zc);
  return (long)T.nx * T.ny * T.nz;
}

int select_shape(int kernel, int device, const nebo_limits_t& lim)
{
  const int pinned = g_pinned[kernel].load();
  if (pinned >= 0) return pinned;
  const int n[3] = {lim.hi[0] - lim.lo[0], lim.hi[1] - lim.lo[1], lim.hi[2] - lim.lo[2]};
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    for (const TunedTile& t : g_tuned)
      if (t.kernel == kernel && t.device == device && t.n[0] == n[0] && t.n[1] == n[1] && t.n[2] == n[2])
        return t.shape;
  }
  const int cus = device_cu_count(device);
  for (int sh : kLongChunks)
    if (tile_groups(sh, lim) >= cus) return sh;
  if (2 * tile_groups(kDefaultShape, lim) >= cus)
    return kernel == 0 && n[0] % kShapes[kDefaultPairShape].tx == 0 ? kDefaultPairShape : kDefaultShape;
  for (int sh : kSmallColumn)
    if (tile_groups(sh, lim) >= 2 * cus) return sh;
  return kSmallColumn[2];
}

int find_shape(int tx, int ty, int zc, int variant)
{
  for (int s = 0; s < kNumShapes; ++s)
    if (kShapes[s].tx == tx && kShapes[s].ty == ty && kShapes[s].zc == zc && kShapes[s].variant == variant)
      return s;
  return -1;
}

// The column schedule addresses a plane through a buffer range of
// (XY + 2X + 2)*8 bytes, which must stay below 2^31; planes that large run the
// direct schedule instead.  Pair tiles need a box width that is a multiple of
// their width (every lane's two columns inside the box); other boxes run the
// default tile.  (kernel 1, the Laplacian, runs pair tiles on its one-column
// schedule, which covers any width.)
int usable_shape(int shape, int X, int XY, int width, int kernel)
{
  if (kernel == 0 && pair_columns(kShapes[shape]) && width % kShapes[shape].tx != 0) shape = kDefaultShape;
  if (kShapes[shape].variant >= 3 && ((long)XY + 2L * X + 2) * 8 >= (1L << 31)) return kFallbackShape;
  return shape;
}

// "div(k grad phi) kernel, tile 128x8x64, column schedule, 4 rows per lane, 2x2
// waves, loads 3 planes ahead" (built once per kernel and tile: launches of a
// time loop only copy it)
std::string tile_backend_text(const char* kernel, int kind, int shape)
{
  const TileShape sh = kShapes[shape];
  std::ostringstream o;
  o << kernel << ", tile " << sh.tx << "x" << sh.ty << "x" << sh.zc;
  if (sh.variant == 0) o << ", direct loads";
  else if (sh.variant <= 2) o << ", LDS-staged rows, " << sh.variant << " planes of prefetch";
  else if (pair_columns(sh) && kind != 2)
    o << ", pair columns (16-byte accesses), " << rows_per_lane(sh) << " rows per lane, " << sh.tx / (2 * kTx)
      << "x" << stacked_waves(sh) << " waves, loads " << prefetch_planes(sh) << " planes ahead";
  else
    o << ", column schedule, " << rows_per_lane(sh) << " rows per lane, " << sh.tx / kTx << "x"
      << stacked_waves(sh) << " waves, loads " << prefetch_planes(sh) << " planes ahead"
      << (store_policy(sh) == 0 ? ", write-back stores" : "");
  return o.str();
}

const std::string& tile_backend(int kind, int shape)
{
  static const char* names[3] = {"div(k grad phi) kernel", "diffusion Euler step kernel",
                                 "Laplacian + Euler kernel"};
  static const std::vector<std::string>* table = [] {
    auto* t = new std::vector<std::string>;
    for (int k = 0; k < 3; ++k)
      for (int sh = 0; sh < kNumShapes; ++sh) t->push_back(tile_backend_text(names[k], k, sh));
    return t;
  }();
  return (*table)[kind * kNumShapes + shape];
}

int launch_divk_shape(int shape, int form, const nebo_field_t* r, const nebo_field_t* k,
                      const nebo_field_t* phi, const nebo_axis_coefs_t& c, double dt,
                      const nebo_limits_t& lim, void* stream)
{
  shape = usable_shape(shape, r->glob[0], r->glob[0] * r->glob[1], lim.hi[0] - lim.lo[0], 0);
  const TileShape sh = kShapes[shape];
  const Geometry geo = tile_geometry(sh, 0);
  Tiles T = make_tiles(lim, sh.tx, geo.tile_y, sh.zc);
  const int nb = T.nx * T.ny * T.nz;
  const int X = r->glob[0], XY = r->glob[0] * r->glob[1];
  BackendScope bs(tile_backend(form == 2 ? 1 : 0, shape));
  const auto fn = divk_table(form, shape);
  const unsigned threads = geo.threads;
  double *rp = frame_origin(r), *kp = frame_origin(k), *pp = frame_origin(phi);
  return emit(stream, CallPtrs{}, [=](void* s, const CallPtrs&) {
    hipLaunchKernelGGL(fn, dim3(nb), dim3(threads), 0, (hipStream_t)s, rp, kp, pp, T, X, XY, c, dt);
    return check_hip(hipGetLastError(), "Nebo Assignment");
  });
}

int launch_lapl_shape(int shape, const nebo_field_t* out, const nebo_field_t* phi,
                      const nebo_axis_coefs_t& c, double sc, const nebo_limits_t& lim, void* stream)
{
  shape = usable_shape(shape, out->glob[0], out->glob[0] * out->glob[1], lim.hi[0] - lim.lo[0], 1);
  const TileShape sh = kShapes[shape];
  const Geometry geo = tile_geometry(sh, 1);
  Tiles T = make_tiles(lim, sh.tx, geo.tile_y, sh.zc);
  const int nb = T.nx * T.ny * T.nz;
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  BackendScope bs(tile_backend(2, shape));
  const auto fn = lapl_table(shape);
  const unsigned threads = geo.threads;
  double *op = frame_origin(out), *pp = frame_origin(phi);
  return emit(stream, CallPtrs{}, [=](void* s, const CallPtrs&) {
    hipLaunchKernelGGL(fn, dim3(nb), dim3(threads), 0, (hipStream_t)s, op, pp, T, X, XY, c, sc);
    return check_hip(hipGetLastError(), "Nebo Assignment");
  });
}

// Periodic fill of the ghost slab of axis d: every ghost cell takes the value of
// the interior cell whose coordinates are its own, wrapped into the interior on
// every axis (so edges and corners are consistent without ordering).
__global__ void fill_periodic_kernel(double* __restrict__ v, int d, int X, int Y,
                                     int o0, int o1, int o2,   // window offset
                                     int e0, int e1, int e2,   // window extent
                                     int g0, int g1, int g2,   // ghost minus
                                     int n0, int n1, int n2,   // interior counts
                                     int gm_d, int gp_d, long total)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  // slab of axis d: (gm_d + gp_d) layers along d, full extent along the others
  int e[3] = {e0, e1, e2};
  e[d] = gm_d + gp_d;
  int c[3];
  c[0] = (int)(t % e[0]);
  long r = t / e[0];
  c[1] = (int)(r % e[1]);
  c[2] = (int)(r / e[1]);
  // map the slab's d coordinate to window coordinates: minus ghosts then plus ghosts
  const int nd = d == 0 ? n0 : d == 1 ? n1 : n2;
  c[d] = c[d] < gm_d ? c[d] : c[d] + nd;
  const int g[3] = {g0, g1, g2};
  const int n[3] = {n0, n1, n2};
  int dst[3], src[3];
  const int o[3] = {o0, o1, o2};
  for (int a = 0; a < 3; ++a) {
    dst[a] = o[a] + c[a];
    int rel = c[a] - g[a];                 // interior-relative coordinate
    if (n[a] > 0) {
      rel %= n[a];
      if (rel < 0) rel += n[a];
    }
    src[a] = o[a] + g[a] + rel;
  }
  v[dst[0] + (long)X * (dst[1] + (long)Y * dst[2])] = v[src[0] + (long)X * (src[1] + (long)Y * src[2])];
}

}  // namespace

int launch_divkgrad(const nebo_field_t* r, const nebo_field_t* k, const nebo_field_t* phi,
                    const nebo_axis_coefs_t& c, int form, const nebo_limits_t& lim, void* stream)
{
  return launch_divk_shape(select_shape(0, r->device, lim), form, r, k, phi, c, 0.0, lim, stream);
}

int launch_diffusion_euler(const nebo_field_t* out, const nebo_field_t* k, const nebo_field_t* phi,
                           const nebo_axis_coefs_t& c, double dt, const nebo_limits_t& lim,
                           void* stream)
{
  return launch_divk_shape(select_shape(0, out->device, lim), 2, out, k, phi, c, dt, lim, stream);
}

int launch_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                      const nebo_axis_coefs_t& c, double sc, const nebo_limits_t& lim,
                      void* stream)
{
  return launch_lapl_shape(select_shape(1, out->device, lim), out, phi, c, sc, lim, stream);
}

int fill_ghosts_periodic(const nebo_field_t* f, void* stream)
{
  static const char* ctx = "Nebo Ghost Fill";
  int n[3];
  for (int d = 0; d < 3; ++d) {
    n[d] = f->extent[d] - f->ghost_minus[d] - f->ghost_plus[d];
    if ((f->ghost_minus[d] > 0 || f->ghost_plus[d] > 0) &&
        (n[d] < f->ghost_minus[d] || n[d] < f->ghost_plus[d]))
      return set_error(NEBO_ERR_GHOST, ctx, "fewer interior cells than ghost layers on a periodic axis");
  }
  for (int d = 0; d < 3; ++d) {
    const int gm = f->ghost_minus[d], gp = f->ghost_plus[d];
    if (gm == 0 && gp == 0) continue;
    long total = gm + gp;
    for (int a = 0; a < 3; ++a)
      if (a != d) total *= f->extent[a];
    const int blk = 256;
    hipLaunchKernelGGL(fill_periodic_kernel, dim3((unsigned)((total + blk - 1) / blk)), dim3(blk), 0,
                       (hipStream_t)stream, static_cast<double*>(f->values), d, f->glob[0], f->glob[1],
                       f->offset[0], f->offset[1], f->offset[2], f->extent[0], f->extent[1],
                       f->extent[2], f->ghost_minus[0], f->ghost_minus[1], f->ghost_minus[2], n[0],
                       n[1], n[2], gm, gp, total);
    int rc = check_hip(hipGetLastError(), ctx);
    if (rc) return rc;
  }
  return NEBO_OK;
}

namespace {

int check_same_shape(const nebo_field_t* a, const nebo_field_t* b, const char* ctx)
{
  for (int d = 0; d < 3; ++d)
    if (a->glob[d] != b->glob[d] || a->extent[d] != b->extent[d] ||
        a->valid_minus[d] != b->valid_minus[d] || a->valid_plus[d] != b->valid_plus[d])
      return set_error(NEBO_ERR_INVALID, ctx, "operand fields must share one memory window and ghost layout");
  if (a->device != b->device || a->device < 0)
    return set_error(NEBO_ERR_DEVICE, ctx, "operands must be resident on one GPU");
  return NEBO_OK;
}

// Cells of f's interior whose one-cell stencil reach stays inside the valid
// ghosts: [max(0, 1 - valid_minus), ext - max(0, 1 - valid_plus)).  *empty when
// no cell qualifies (e.g. an axis of extent 1 without ghosts).
void stencil_range(const nebo_field_t* f, nebo_limits_t* L, int* empty)
{
  *empty = 0;
  for (int d = 0; d < 3; ++d) {
    const int ext = f->extent[d] - f->valid_minus[d] - f->valid_plus[d];
    L->lo[d] = f->valid_minus[d] >= 1 ? 0 : 1 - f->valid_minus[d];
    L->hi[d] = ext - (f->valid_plus[d] >= 1 ? 0 : 1 - f->valid_plus[d]);
    if (L->hi[d] <= L->lo[d]) *empty = 1;
  }
}

}  // namespace

namespace {
constexpr int kMaxDevices = 64;
std::atomic<int> g_cus[kMaxDevices];
}  // namespace

int device_cu_count(int device)
{
  if (device >= 0 && device < kMaxDevices && g_cus[device].load() > 0) return g_cus[device].load();
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
  if (device >= 0 && device < kMaxDevices) g_cus[device].store(n);
  return n;
}

}  // namespace nebo

using namespace nebo;

extern "C" int nebo_fill_ghosts_periodic(const nebo_field_t* field, void* stream)
{
  clear_error();
  if (!field) return set_error(NEBO_ERR_INVALID, "Nebo Ghost Fill", "null argument");
  if (field->device < 0) return set_error(NEBO_ERR_DEVICE, "Nebo Ghost Fill", "field is not on a GPU");
  return fill_ghosts_periodic(field, stream);
}

extern "C" int nebo_divkgrad(const nebo_field_t* r, const nebo_field_t* k,
                             const nebo_field_t* phi, const nebo_axis_coefs_t* c, int form,
                             void* stream)
{
  static const char* ctx = "Nebo Assignment";
  clear_error();
  if (!r || !k || !phi || !c || (form != 0 && form != 1))
    return set_error(NEBO_ERR_INVALID, ctx, "null argument or unknown form");
  int rc;
  if ((rc = check_same_shape(r, k, ctx)) || (rc = check_same_shape(r, phi, ctx))) return rc;
  if (r->values == k->values || r->values == phi->values)
    return set_error(NEBO_ERR_INVALID, ctx, "result must not alias an operand");
  nebo_limits_t L;
  int empty = 0;
  stencil_range(k, &L, &empty);
  if (empty) return NEBO_OK;
  return launch_divkgrad(r, k, phi, *c, form, L, stream);
}

extern "C" int nebo_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                               const nebo_axis_coefs_t* c, double s, void* stream)
{
  static const char* ctx = "Nebo Assignment";
  clear_error();
  if (!out || !phi || !c) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  int rc;
  if ((rc = check_same_shape(out, phi, ctx))) return rc;
  if (out->values == phi->values) return set_error(NEBO_ERR_INVALID, ctx, "result must not alias the operand");
  nebo_limits_t L;
  int empty = 0;
  stencil_range(phi, &L, &empty);
  if (empty) return NEBO_OK;
  return launch_lapl_euler(out, phi, *c, s, L, stream);
}

extern "C" int nebo_stencil_set_tile(int kernel, int32_t tx, int32_t ty, int32_t zchunk, int32_t variant)
{
  clear_error();
  if ((kernel == 0 || kernel == 1) && tx == 0 && ty == 0 && zchunk == 0 && variant == 0) {
    // unpin: tuned tiles and the size rule apply again
    g_pinned[kernel].store(-1);
    g_last[kernel].store(kDefaultShape);
    bump_plan_epoch();
    return NEBO_OK;
  }
  const int s = find_shape(tx, ty, zchunk, variant);
  if (s < 0 || (kernel != 0 && kernel != 1)) {
    std::ostringstream m;
    m << "no stencil kernel instantiated for kernel " << kernel << ", tile " << tx << "x" << ty << "x" << zchunk
      << ", variant " << variant;
    return set_error(NEBO_ERR_INVALID, "Nebo Stencil Tile", m.str());
  }
  g_pinned[kernel].store(s);
  g_last[kernel].store(s);
  bump_plan_epoch();
  return NEBO_OK;
}

extern "C" int nebo_stencil_get_tile(int kernel, int32_t* tx, int32_t* ty, int32_t* zchunk, int32_t* variant)
{
  clear_error();
  if (kernel != 0 && kernel != 1) return set_error(NEBO_ERR_INVALID, "Nebo Stencil Tile", "unknown kernel");
  const TileShape sh = kShapes[g_last[kernel].load()];
  if (tx) *tx = sh.tx;
  if (ty) *ty = sh.ty;
  if (zchunk) *zchunk = sh.zc;
  if (variant) *variant = sh.variant;
  return NEBO_OK;
}

extern "C" int nebo_stencil_select_tile(int kernel, int device, int32_t nx, int32_t ny, int32_t nz, int32_t* tx,
                                        int32_t* ty, int32_t* zchunk, int32_t* variant)
{
  clear_error();
  if ((kernel != 0 && kernel != 1) || nx < 1 || ny < 1 || nz < 1 || device < 0)
    return set_error(NEBO_ERR_INVALID, "Nebo Stencil Tile", "unknown kernel, bad device or empty box");
  nebo_limits_t L;
  const int32_t n[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) { L.lo[d] = 0; L.hi[d] = n[d]; }
  const TileShape sh = kShapes[select_shape(kernel, device, L)];
  if (tx) *tx = sh.tx;
  if (ty) *ty = sh.ty;
  if (zchunk) *zchunk = sh.zc;
  if (variant) *variant = sh.variant;
  return NEBO_OK;
}

extern "C" int nebo_stencil_autotune(int kernel, const nebo_field_t* r, const nebo_field_t* k,
                                     const nebo_field_t* phi, const nebo_axis_coefs_t* c,
                                     void* stream, int32_t* tx, int32_t* ty, int32_t* zchunk,
                                     int32_t* variant, float* times_ms)
{
  static const char* ctx = "Nebo Stencil Tile";
  clear_error();
  if (kernel != 0 && kernel != 1) return set_error(NEBO_ERR_INVALID, ctx, "unknown kernel");
  if (!r || !phi || !c || (kernel == 0 && !k)) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  int rc;
  if ((rc = check_same_shape(r, phi, ctx))) return rc;
  if (kernel == 0 && (rc = check_same_shape(r, k, ctx))) return rc;
  if (r->values == phi->values || (kernel == 0 && r->values == k->values))
    return set_error(NEBO_ERR_INVALID, ctx, "result must not alias an operand");
  nebo_limits_t L;
  int empty = 0;
  stencil_range(phi, &L, &empty);
  if (empty) return set_error(NEBO_ERR_INVALID, ctx, "no interior cells to time");
  auto run = [&](int sh) {
    return kernel == 0 ? launch_divk_shape(sh, 0, r, k, phi, *c, 0.0, L, stream)
                       : launch_lapl_shape(sh, r, phi, *c, 0.5, L, stream);
  };
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  if ((rc = check_hip(hipEventCreate(&e0), ctx))) return rc;
  if ((rc = check_hip(hipEventCreate(&e1), ctx))) {
    (void)hipEventDestroy(e0);
    return rc;
  }
  int best = kDefaultShape;
  float best_ms = 1e30f;
  for (int sh = 0; sh < kNumShapes && !rc; ++sh) {
    rc = run(sh);   // warm
    if (!rc) rc = check_hip(hipEventRecord(e0, s), ctx);
    for (int rep = 0; rep < 5 && !rc; ++rep) rc = run(sh);
    if (!rc) rc = check_hip(hipEventRecord(e1, s), ctx);
    if (!rc) rc = check_hip(hipEventSynchronize(e1), ctx);
    float ms = 0.f;
    if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), ctx);
    ms /= 5.0f;
    if (times_ms) times_ms[sh] = ms;
    if (!rc && ms < best_ms) {
      best_ms = ms;
      best = sh;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  {
    // remembered for boxes of these dimensions on this device
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    const int n[3] = {L.hi[0] - L.lo[0], L.hi[1] - L.lo[1], L.hi[2] - L.lo[2]};
    bool found = false;
    for (TunedTile& t : g_tuned)
      if (t.kernel == kernel && t.device == r->device && t.n[0] == n[0] && t.n[1] == n[1] && t.n[2] == n[2]) {
        t.shape = best;
        found = true;
      }
    if (!found) g_tuned.push_back(TunedTile{kernel, r->device, {n[0], n[1], n[2]}, best});
  }
  g_pinned[kernel].store(-1);   // the tuned tile takes effect for its box size
  g_last[kernel].store(best);
  bump_plan_epoch();
  if (tx) *tx = kShapes[best].tx;
  if (ty) *ty = kShapes[best].ty;
  if (zchunk) *zchunk = kShapes[best].zc;
  if (variant) *variant = kShapes[best].variant;
  return NEBO_OK;
}

extern "C" int nebo_stencil_tiles(int32_t* n, int32_t* tx, int32_t* ty, int32_t* zchunk, int32_t* variant)
{
  clear_error();
  if (n) *n = kNumShapes;
  for (int s = 0; s < kNumShapes; ++s) {
    if (tx) tx[s] = kShapes[s].tx;
    if (ty) ty[s] = kShapes[s].ty;
    if (zchunk) zchunk[s] = kShapes[s].zc;
    if (variant) variant[s] = kShapes[s].variant;
  }
  return NEBO_OK;
}
