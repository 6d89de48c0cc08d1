// nebo_transport.hip -- scalar transport chain: one SSP-RK3 stage of
//   R(s) = div(interp(k)*grad(s)) - div(u*upwind(s)),
//   upwind(s) on face f = cond(u_f > 0, s[cell below f])(s[cell above f]),
// its transpose, and the checkpointed forward / reverse sweeps over many steps.
//
// The forward stage evaluates each cell in the order the fused expression
//   ((Dx(Ix(k)*Gx(s)) + Dy(..)) + Dz(..)) - ((Dx(c_x*u_x) + Dy(..)) + Dz(..))
// is evaluated by the generic evaluator (stencil folds of NeboStencils.h:407-448,
// the first-true-clause select of NeboCond.h:408-412), so it is bit-identical to
// nebo_assign of that stage expression.  Periodicity enters through one ghost
// layer holding the wrapped interior; every stage refills its output's ghosts.
//
// Transpose of a stage, for adjoint a and w_d = D_d^T a on the faces of axis d:
//   R_s^T a    = sum_d G_d^T(kf_d .* w_d) - sum_d U_d^T(u_d .* w_d)
//   R_k(s)^T a = sum_d I_d^T(w_d .* G_d s)
//   (D^T a)_f = a_f*d0 + a_{f-1}*d1,  (G^T v)_c = v_{c+1}*g0 + v_c*g1,
//   (I^T v)_c = v_{c+1}*i0 + v_c*i1,  (U^T v)_c = [u_{c+1} > 0] v_{c+1} + [!(u_c > 0)] v_c
// With X = a + dt*R_s^T a:
//   stage 3: out = (2/3)*X,  acc_out = (1/3)*a,          gk += ((2/3)*dt) * R_k(s2)^T a
//   stage 2: out = 0.25*X,   acc_out = acc_in + 0.75*a,  gk += (0.25*dt) * R_k(s1)^T a
//   stage 1: out = acc_in + X,                            gk += dt * R_k(phi_n)^T a
// One fused pass per adjoint stage; z faces are carried in registers.

#include "nebo_internal.h"
#include "nebo_device.h"

#include <hip/hip_runtime.h>

#include <atomic>

namespace nebo {
namespace {

constexpr int kTx = 64;
constexpr int kTy = 4;
constexpr int kZChunk = 16;
const char* kCtx = "Nebo Transport";

struct AxisC {
  double i0, i1, g0, g1, d0, d1;
};
struct Coefs3 {
  AxisC a[3];
};

Coefs3 coefs3(const nebo_axis_coefs_t& c)
{
  Coefs3 r;
  for (int d = 0; d < 3; ++d)
    r.a[d] = AxisC{c.interp[d][0], c.interp[d][1], c.grad[d][0], c.grad[d][1], c.div[d][0], c.div[d][1]};
  return r;
}

// upwinded convective flux on one face: cond(u > 0, s_lo)(s_hi) * u
__device__ __forceinline__ double conv_face(double u, double s_lo, double s_hi)
{
  return (u > 0.0 ? s_lo : s_hi) * u;
}

template <int STAGE>
__global__ __launch_bounds__(kTx * kTy) void transport_stage_kernel(
    double* __restrict__ out, const double* __restrict__ pn, const double* __restrict__ s,
    const double* __restrict__ k, const double* __restrict__ ux, const double* __restrict__ uy,
    const double* __restrict__ uz, Box3 B, int X, int XY, Coefs3 C, double dt)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % B.nx;
  const int rest = t / B.nx;
  const int ty = rest % B.ny;
  const int tz = rest / B.ny;
  const int x = tx * kTx + (threadIdx.x & (kTx - 1));
  const int y = ty * kTy + (threadIdx.x / kTx);
  if (x >= B.n[0] || y >= B.n[1]) return;
  const int z0 = tz * B.zc;
  const int z1 = min(z0 + B.zc, B.n[2]);
  const AxisC cx = C.a[0], cy = C.a[1], cz = C.a[2];

  int i = x + X * y + XY * z0;
  double sc = s[i], kc = k[i];
  double fz_lo, vz_lo;
  {
    const double szm = s[i - XY], kzm = k[i - XY], u = uz[i];
    fz_lo = face_flux(kzm, kc, szm, sc, cz.i0, cz.i1, cz.g0, cz.g1);
    vz_lo = conv_face(u, szm, sc);
  }
  for (int z = z0; z < z1; ++z, i += XY) {
    const double szp = s[i + XY], kzp = k[i + XY], uzh = uz[i + XY];
    const double sxm = s[i - 1], sxp = s[i + 1], kxm = k[i - 1], kxp = k[i + 1];
    const double sym = s[i - X], syp = s[i + X], kym = k[i - X], kyp = k[i + X];
    const double uxl = ux[i], uxh = ux[i + 1], uyl = uy[i], uyh = uy[i + X];
    const double fx = face_flux(kxm, kc, sxm, sc, cx.i0, cx.i1, cx.g0, cx.g1) * cx.d0 +
                      face_flux(kc, kxp, sc, sxp, cx.i0, cx.i1, cx.g0, cx.g1) * cx.d1;
    const double fy = face_flux(kym, kc, sym, sc, cy.i0, cy.i1, cy.g0, cy.g1) * cy.d0 +
                      face_flux(kc, kyp, sc, syp, cy.i0, cy.i1, cy.g0, cy.g1) * cy.d1;
    const double fz_hi = face_flux(kc, kzp, sc, szp, cz.i0, cz.i1, cz.g0, cz.g1);
    const double fz = fz_lo * cz.d0 + fz_hi * cz.d1;
    const double vx = conv_face(uxl, sxm, sc) * cx.d0 + conv_face(uxh, sc, sxp) * cx.d1;
    const double vy = conv_face(uyl, sym, sc) * cy.d0 + conv_face(uyh, sc, syp) * cy.d1;
    const double vz_hi = conv_face(uzh, sc, szp);
    const double vz = vz_lo * cz.d0 + vz_hi * cz.d1;
    const double R = ((fx + fy) + fz) - ((vx + vy) + vz);
    double v;
    if constexpr (STAGE == 1) v = pn[i] + dt * R;
    else if constexpr (STAGE == 2) v = 0.75 * pn[i] + 0.25 * (sc + dt * R);
    else v = (1.0 / 3.0) * pn[i] + (2.0 / 3.0) * (sc + dt * R);
    out[i] = v;
    sc = szp;
    kc = kzp;
    fz_lo = fz_hi;
    vz_lo = vz_hi;
  }
}

// Column schedule of the stage (the layout of divkgrad_column_kernel): W waves
// side by side in x and NW stacked in y per workgroup, each lane one x column of
// RY rows marching zc planes.  Per plane a lane holds s and k on its rows plus
// one halo row on each side, u_x on its rows (the high x face is the next lane's
// low one), u_y on its rows and the row above, u_z on its rows (plane z+1's
// values are the high z faces) and phi_n; x neighbours come from the adjacent
// lanes by DPP and, at the wave ends, from the edge loads (readlane), y
// neighbours from the lane's own rows.  Loads run PF planes ahead in PF+1
// rotated register slots.  Every cell is evaluated in transport_stage_kernel's
// order, so the two are bit-identical.  Offsets are 32-bit buffer offsets from
// each plane's base (launcher: plane sizes below 2^31 bytes).  A plane slot is
// 32 doubles, so the kernel takes the register file of one wave per SIMD (two
// with loads one plane ahead): one 4-wave workgroup per CU, no spills.
template <int STAGE, int W, int NW, int RY, int PF>
__global__ __launch_bounds__(kLanes * W * NW) __attribute__((amdgpu_waves_per_eu(1, PF >= 2 ? 1 : 2)))
void transport_column_kernel(
    double* __restrict__ out, const double* __restrict__ pn, const double* __restrict__ s,
    const double* __restrict__ k, const double* __restrict__ ux, const double* __restrict__ uy,
    const double* __restrict__ uz, Tiles T, int zc, int X, int XY, Coefs3 C, double dt)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = threadIdx.x / kLanes;
  const int xw0 = T.lo[0] + tx * (kLanes * W) + (wave % W) * kLanes;
  const int x = xw0 + lane;
  const int y0 = T.lo[1] + ty * (RY * NW) + (wave / W) * RY;
  const int z0 = T.lo[2] + tz * zc;
  const int z1 = min(z0 + zc, T.hi[2]);
  const bool xin = x < T.hi[0];
  const ColumnGeom<RY> g = column_geom<RY>(T, X, xw0, y0, lane);
  const AxisC cx = C.a[0], cy = C.a[1], cz = C.a[2];

  struct Col {
    double s[RY + 2], k[RY + 2], se, ke, ux[RY], uxe, uy[RY + 1], uz[RY], pn[RY];
  };
  auto load = [&](int z) __attribute__((always_inline)) {
    Col v;
    const Rsrc sb = plane_rsrc(plane_base(s, T, X, XY, z), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z), X, XY);
    const Rsrc xb = plane_rsrc(plane_base(ux, T, X, XY, z), X, XY);
    const Rsrc yb = plane_rsrc(plane_base(uy, T, X, XY, z), X, XY);
    const Rsrc zb = plane_rsrc(plane_base(uz, T, X, XY, z), X, XY);
#pragma unroll
    for (int q = 0; q < RY + 2; ++q) {
      v.s[q] = load_at(sb, g.row[q]);
      v.k[q] = load_at(kb, g.row[q]);
    }
    v.se = load_at(sb, g.edge);
    v.ke = load_at(kb, g.edge);
    v.uxe = load_at(xb, g.edge);
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      v.ux[j] = load_at(xb, g.row[j + 1]);
      v.uz[j] = load_at(zb, g.row[j + 1]);
    }
#pragma unroll
    for (int j = 0; j < RY + 1; ++j) v.uy[j] = load_at(yb, g.row[j + 1]);
    if constexpr (STAGE != 1) {
      const Rsrc nb = plane_rsrc(plane_base(pn, T, X, XY, z), X, XY);
#pragma unroll
      for (int j = 0; j < RY; ++j) v.pn[j] = load_at(nb, g.row[j + 1]);
    } else {
#pragma unroll
      for (int j = 0; j < RY; ++j) v.pn[j] = 0.0;   // stage 1 reads phi_n through s
    }
    return v;
  };

  constexpr int NS = PF + 1;   // register slots: planes z .. z+PF
  Col S[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) S[q] = load(min(z0 + q, z1));
  // low z faces of plane z0
  double fz[RY], vz[RY];
  {
    const Rsrc sb = plane_rsrc(plane_base(s, T, X, XY, z0 - 1), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z0 - 1), X, XY);
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const double szm = load_at(sb, g.row[j + 1]), kzm = load_at(kb, g.row[j + 1]);
      const double sc = S[0].s[j + 1], kc = S[0].k[j + 1];
      fz[j] = face_flux(kzm, kc, szm, sc, cz.i0, cz.i1, cz.g0, cz.g1);
      vz[j] = conv_face(S[0].uz[j], szm, sc);
    }
  }
  // one step: plane z, whose slot is jj (a compile-time index)
  auto step = [&](auto JJ, const int z) __attribute__((always_inline)) {
    constexpr int jj = decltype(JJ)::value;
    {
      const Col& C0 = S[jj];              // plane z
      const Col& C1 = S[(jj + 1) % NS];   // plane z+1
      const Rsrc ob = plane_rsrc(plane_base(out, T, X, XY, z), X, XY);
#pragma unroll
      for (int j = 0; j < RY; ++j) {
        const double sc = C0.s[j + 1], kc = C0.k[j + 1];
        const double sl = from_lower_lane(sc), su = from_upper_lane(sc);
        const double kl = from_lower_lane(kc), ku = from_upper_lane(kc);
        const double uu = from_upper_lane(C0.ux[j]);
        // The edge lanes read inside the selects: this kernel keeps its plane
        // slots in VGPRs (213-237, no AGPRs on gfx950), so lane j's copy of an
        // edge value is current when the edge lane reads it; reading them in
        // every lane first (as lane_value asks where values may be spilled)
        // measured stage 1 at 1.37 ms against 1.26 at 512^3.
        const double sxm = lane == 0 ? lane_value(C0.se, j) : sl;
        const double sxp = lane == kLanes - 1 ? lane_value(C0.se, 32 + j) : su;
        const double kxm = lane == 0 ? lane_value(C0.ke, j) : kl;
        const double kxp = lane == kLanes - 1 ? lane_value(C0.ke, 32 + j) : ku;
        const double uxh = lane == kLanes - 1 ? lane_value(C0.uxe, 32 + j) : uu;
        const double sym = C0.s[j], syp = C0.s[j + 2], kym = C0.k[j], kyp = C0.k[j + 2];
        const double szp = C1.s[j + 1], kzp = C1.k[j + 1];
        const double fx = face_flux(kxm, kc, sxm, sc, cx.i0, cx.i1, cx.g0, cx.g1) * cx.d0 +
                          face_flux(kc, kxp, sc, sxp, cx.i0, cx.i1, cx.g0, cx.g1) * cx.d1;
        const double fy = face_flux(kym, kc, sym, sc, cy.i0, cy.i1, cy.g0, cy.g1) * cy.d0 +
                          face_flux(kc, kyp, sc, syp, cy.i0, cy.i1, cy.g0, cy.g1) * cy.d1;
        const double fz_hi = face_flux(kc, kzp, sc, szp, cz.i0, cz.i1, cz.g0, cz.g1);
        const double fzz = fz[j] * cz.d0 + fz_hi * cz.d1;
        const double vx = conv_face(C0.ux[j], sxm, sc) * cx.d0 + conv_face(uxh, sc, sxp) * cx.d1;
        const double vy = conv_face(C0.uy[j], sym, sc) * cy.d0 + conv_face(C0.uy[j + 1], sc, syp) * cy.d1;
        const double vz_hi = conv_face(C1.uz[j], sc, szp);
        const double vzz = vz[j] * cz.d0 + vz_hi * cz.d1;
        const double R = ((fx + fy) + fzz) - ((vx + vy) + vzz);
        double v;
        if constexpr (STAGE == 1) v = sc + dt * R;   // phi_n is s
        else if constexpr (STAGE == 2) v = 0.75 * C0.pn[j] + 0.25 * (sc + dt * R);
        else v = (1.0 / 3.0) * C0.pn[j] + (2.0 / 3.0) * (sc + dt * R);
        fz[j] = fz_hi;
        vz[j] = vz_hi;
        if (xin && y0 + j < T.hi[1]) store_nt_at(ob, g.row[j + 1], v);
      }
      S[jj] = load(min(z + NS, z1));   // plane z is consumed: its slot takes z+NS
    }
  };
  const int full = (z1 - z0) / NS, rem = (z1 - z0) - NS * full;
  int z = z0;
  for (int q = 0; q < full; ++q, z += NS)
    unroll_steps([&](auto J) { step(J, z + decltype(J)::value); }, std::make_integer_sequence<int, NS>{});
  unroll_steps([&](auto J) { if (rem > decltype(J)::value) step(J, z + decltype(J)::value); },
               std::make_integer_sequence<int, NS - 1>{});
}

// Adjoint quantities of one face: kf*w, w*(G s), u*w and u itself.
struct AdjFace {
  double kw, wg, uw, u;
};

__device__ __forceinline__ AdjFace adj_face(double a_lo, double a_hi, double s_lo, double s_hi,
                                            double k_lo, double k_hi, double u, const AxisC& c)
{
  const double w = a_hi * c.d0 + a_lo * c.d1;
  const double gs = s_lo * c.g0 + s_hi * c.g1;
  const double kf = k_lo * c.i0 + k_hi * c.i1;
  return AdjFace{kf * w, w * gs, u * w, u};
}

struct AdjCell {
  double A, V, B;
};

__device__ __forceinline__ AdjCell adj_cell(const AdjFace& lo, const AdjFace& hi, const AxisC& c)
{
  return AdjCell{hi.kw * c.g0 + lo.kw * c.g1, (hi.u > 0.0 ? hi.uw : 0.0) + (lo.u > 0.0 ? 0.0 : lo.uw),
                 hi.wg * c.i0 + lo.wg * c.i1};
}

template <int STAGE>
__global__ __launch_bounds__(kTx * kTy) void transport_adjoint_kernel(
    double* __restrict__ out, double* __restrict__ acc_out, double* __restrict__ gk,
    const double* __restrict__ a, const double* __restrict__ acc_in, const double* __restrict__ s,
    const double* __restrict__ k, const double* __restrict__ ux, const double* __restrict__ uy,
    const double* __restrict__ uz, Box3 B, int X, int XY, Coefs3 C, double dt, double cs, double cdt)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % B.nx;
  const int rest = t / B.nx;
  const int ty = rest % B.ny;
  const int tz = rest / B.ny;
  const int x = tx * kTx + (threadIdx.x & (kTx - 1));
  const int y = ty * kTy + (threadIdx.x / kTx);
  if (x >= B.n[0] || y >= B.n[1]) return;
  const int z0 = tz * B.zc;
  const int z1 = min(z0 + B.zc, B.n[2]);
  const AxisC cx = C.a[0], cy = C.a[1], cz = C.a[2];

  int i = x + X * y + XY * z0;
  double ac = a[i], sc = s[i], kc = k[i];
  AdjFace zlo = adj_face(a[i - XY], ac, s[i - XY], sc, k[i - XY], kc, uz[i], cz);
  for (int z = z0; z < z1; ++z, i += XY) {
    const double az = a[i + XY], szp = s[i + XY], kzp = k[i + XY];
    const AdjFace xlo = adj_face(a[i - 1], ac, s[i - 1], sc, k[i - 1], kc, ux[i], cx);
    const AdjFace xhi = adj_face(ac, a[i + 1], sc, s[i + 1], kc, k[i + 1], ux[i + 1], cx);
    const AdjFace ylo = adj_face(a[i - X], ac, s[i - X], sc, k[i - X], kc, uy[i], cy);
    const AdjFace yhi = adj_face(ac, a[i + X], sc, s[i + X], kc, k[i + X], uy[i + X], cy);
    const AdjFace zhi = adj_face(ac, az, sc, szp, kc, kzp, uz[i + XY], cz);
    const AdjCell ex = adj_cell(xlo, xhi, cx), ey = adj_cell(ylo, yhi, cy), ez = adj_cell(zlo, zhi, cz);
    const double rt = ((ex.A + ey.A) + ez.A) - ((ex.V + ey.V) + ez.V);
    const double Xv = ac + dt * rt;
    if constexpr (STAGE == 3) {
      out[i] = cs * Xv;
      acc_out[i] = (1.0 / 3.0) * ac;
    } else if constexpr (STAGE == 2) {
      out[i] = cs * Xv;
      acc_out[i] = acc_in[i] + 0.75 * ac;
    } else {
      out[i] = acc_in[i] + Xv;
    }
    gk[i] = gk[i] + cdt * ((ex.B + ey.B) + ez.B);
    ac = az;
    sc = szp;
    kc = kzp;
    zlo = zhi;
  }
}

// The adjoint stage on column tiles (the layout of transport_column_kernel): a,
// s and k on the lane's rows plus one halo row each side and the wave-end edges,
// the face velocities as in the stage kernel, acc_in and g_k on the lane's rows.
// Every cell is evaluated in transport_adjoint_kernel's order (bit-identical);
// g_k is read before its cell is written and read by no other cell.
template <int STAGE, int W, int NW, int RY, int PF>
__global__ __launch_bounds__(kLanes * W * NW) __attribute__((amdgpu_waves_per_eu(1, PF >= 2 ? 1 : 2)))
void transport_adjoint_column_kernel(
    double* __restrict__ out, double* __restrict__ acc_out, double* __restrict__ gk,
    const double* __restrict__ a, const double* __restrict__ acc_in, const double* __restrict__ s,
    const double* __restrict__ k, const double* __restrict__ ux, const double* __restrict__ uy,
    const double* __restrict__ uz, Tiles T, int zc, int X, int XY, Coefs3 C, double dt, double cs, double cdt)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = threadIdx.x / kLanes;
  const int xw0 = T.lo[0] + tx * (kLanes * W) + (wave % W) * kLanes;
  const int x = xw0 + lane;
  const int y0 = T.lo[1] + ty * (RY * NW) + (wave / W) * RY;
  const int z0 = T.lo[2] + tz * zc;
  const int z1 = min(z0 + zc, T.hi[2]);
  const bool xin = x < T.hi[0];
  const ColumnGeom<RY> g = column_geom<RY>(T, X, xw0, y0, lane);
  const AxisC cx = C.a[0], cy = C.a[1], cz = C.a[2];

  struct Col {
    double a[RY + 2], s[RY + 2], k[RY + 2], ae, se, ke, ux[RY], uxe, uy[RY + 1], uz[RY], ain[RY], gk[RY];
  };
  auto load = [&](int z) __attribute__((always_inline)) {
    Col v;
    const Rsrc ab = plane_rsrc(plane_base(a, T, X, XY, z), X, XY);
    const Rsrc sb = plane_rsrc(plane_base(s, T, X, XY, z), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z), X, XY);
    const Rsrc xb = plane_rsrc(plane_base(ux, T, X, XY, z), X, XY);
    const Rsrc yb = plane_rsrc(plane_base(uy, T, X, XY, z), X, XY);
    const Rsrc zb = plane_rsrc(plane_base(uz, T, X, XY, z), X, XY);
    const Rsrc gb = plane_rsrc(plane_base(static_cast<const double*>(gk), T, X, XY, z), X, XY);
#pragma unroll
    for (int q = 0; q < RY + 2; ++q) {
      v.a[q] = load_at(ab, g.row[q]);
      v.s[q] = load_at(sb, g.row[q]);
      v.k[q] = load_at(kb, g.row[q]);
    }
    v.ae = load_at(ab, g.edge);
    v.se = load_at(sb, g.edge);
    v.ke = load_at(kb, g.edge);
    v.uxe = load_at(xb, g.edge);
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      v.ux[j] = load_at(xb, g.row[j + 1]);
      v.uz[j] = load_at(zb, g.row[j + 1]);
      v.gk[j] = load_at(gb, g.row[j + 1]);
    }
#pragma unroll
    for (int j = 0; j < RY + 1; ++j) v.uy[j] = load_at(yb, g.row[j + 1]);
    if constexpr (STAGE != 3) {
      const Rsrc ib = plane_rsrc(plane_base(acc_in, T, X, XY, z), X, XY);
#pragma unroll
      for (int j = 0; j < RY; ++j) v.ain[j] = load_at(ib, g.row[j + 1]);
    } else {
#pragma unroll
      for (int j = 0; j < RY; ++j) v.ain[j] = 0.0;   // stage 3 reads no acc_in
    }
    return v;
  };

  constexpr int NS = PF + 1;
  Col S[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) S[q] = load(min(z0 + q, z1));
  AdjFace zlo[RY];
  {
    const Rsrc ab = plane_rsrc(plane_base(a, T, X, XY, z0 - 1), X, XY);
    const Rsrc sb = plane_rsrc(plane_base(s, T, X, XY, z0 - 1), X, XY);
    const Rsrc kb = plane_rsrc(plane_base(k, T, X, XY, z0 - 1), X, XY);
#pragma unroll
    for (int j = 0; j < RY; ++j)
      zlo[j] = adj_face(load_at(ab, g.row[j + 1]), S[0].a[j + 1], load_at(sb, g.row[j + 1]), S[0].s[j + 1],
                        load_at(kb, g.row[j + 1]), S[0].k[j + 1], S[0].uz[j], cz);
  }
  auto step = [&](auto JJ, const int z) __attribute__((always_inline)) {
    constexpr int jj = decltype(JJ)::value;
    {
      const Col& C0 = S[jj];
      const Col& C1 = S[(jj + 1) % NS];
      const Rsrc ob = plane_rsrc(plane_base(out, T, X, XY, z), X, XY);
      const Rsrc gb = plane_rsrc(plane_base(gk, T, X, XY, z), X, XY);
#pragma unroll
      for (int j = 0; j < RY; ++j) {
        const double ac = C0.a[j + 1], sc = C0.s[j + 1], kc = C0.k[j + 1];
        const double al = from_lower_lane(ac), au = from_upper_lane(ac);
        const double sl = from_lower_lane(sc), su = from_upper_lane(sc);
        const double kl = from_lower_lane(kc), ku = from_upper_lane(kc);
        const double uu = from_upper_lane(C0.ux[j]);
        const bool first = lane == 0, last = lane == kLanes - 1;
        // lane reads outside the selects: every lane takes them (see lane_value)
        const double ael = lane_value(C0.ae, j), aer = lane_value(C0.ae, 32 + j);
        const double sel = lane_value(C0.se, j), ser = lane_value(C0.se, 32 + j);
        const double kel = lane_value(C0.ke, j), ker = lane_value(C0.ke, 32 + j);
        const double uxe = lane_value(C0.uxe, 32 + j);
        const double axm = first ? ael : al, axp = last ? aer : au;
        const double sxm = first ? sel : sl, sxp = last ? ser : su;
        const double kxm = first ? kel : kl, kxp = last ? ker : ku;
        const double uxh = last ? uxe : uu;
        const AdjFace xlo = adj_face(axm, ac, sxm, sc, kxm, kc, C0.ux[j], cx);
        const AdjFace xhi = adj_face(ac, axp, sc, sxp, kc, kxp, uxh, cx);
        const AdjFace ylo = adj_face(C0.a[j], ac, C0.s[j], sc, C0.k[j], kc, C0.uy[j], cy);
        const AdjFace yhi = adj_face(ac, C0.a[j + 2], sc, C0.s[j + 2], kc, C0.k[j + 2], C0.uy[j + 1], cy);
        const AdjFace zhi = adj_face(ac, C1.a[j + 1], sc, C1.s[j + 1], kc, C1.k[j + 1], C1.uz[j], cz);
        const AdjCell ex = adj_cell(xlo, xhi, cx), ey = adj_cell(ylo, yhi, cy), ez = adj_cell(zlo[j], zhi, cz);
        const double rt = ((ex.A + ey.A) + ez.A) - ((ex.V + ey.V) + ez.V);
        const double Xv = ac + dt * rt;
        const double gnew = C0.gk[j] + cdt * ((ex.B + ey.B) + ez.B);
        zlo[j] = zhi;
        if (xin && y0 + j < T.hi[1]) {
          const uint32_t o = g.row[j + 1];
          if constexpr (STAGE == 3) {
            store_nt_at(ob, o, cs * Xv);
            store_nt_at(plane_rsrc(plane_base(acc_out, T, X, XY, z), X, XY), o, (1.0 / 3.0) * ac);
          } else if constexpr (STAGE == 2) {
            store_nt_at(ob, o, cs * Xv);
            store_nt_at(plane_rsrc(plane_base(acc_out, T, X, XY, z), X, XY), o, C0.ain[j] + 0.75 * ac);
          } else {
            store_nt_at(ob, o, C0.ain[j] + Xv);
          }
          store_nt_at(gb, o, gnew);
        }
      }
      S[jj] = load(min(z + NS, z1));
    }
  };
  const int full = (z1 - z0) / NS, rem = (z1 - z0) - NS * full;
  int z = z0;
  for (int q = 0; q < full; ++q, z += NS)
    unroll_steps([&](auto J) { step(J, z + decltype(J)::value); }, std::make_integer_sequence<int, NS>{});
  unroll_steps([&](auto J) { if (rem > decltype(J)::value) step(J, z + decltype(J)::value); },
               std::make_integer_sequence<int, NS - 1>{});
}

struct Operands {
  const nebo_field_t* u;   // three face fields
  const nebo_field_t* k;
};

// Stage schedule: 0 chosen by box size; 1 direct (64 x 4 columns); 2 column
// tiles with loads 2 planes ahead; 3 column tiles with loads 1 plane ahead.
// g_zchunk: planes per workgroup, 0 for the schedule's default.
std::atomic<int> g_schedule{0};
std::atomic<int> g_zchunk{0};

constexpr int kColW = 2, kColNW = 2, kColRY = 4;   // column tiles of 128 x 8 cells

// Schedule and z chunk of a stage over an n[0] x n[1] x n[2] interior.  By box
// size: column tiles (loads 1 plane ahead) with the longest chunk (512 .. 64
// planes) that gives every CU a workgroup, as the stencil kernels do, else the
// direct schedule.
void stage_schedule(const nebo_field_t* f, int* sched, int* zc)
{
  const int n[3] = {f->extent[0] - 2, f->extent[1] - 2, f->extent[2] - 2};
  *sched = g_schedule.load();
  int auto_zc = 64;
  if (*sched == 0) {
    *sched = 1;
    const long cols = (long)((n[0] + kLanes * kColW - 1) / (kLanes * kColW)) *
                      ((n[1] + kColRY * kColNW - 1) / (kColRY * kColNW));
    const int cus = device_cu_count(f->device);
    for (int c : {512, 256, 128, 64})
      if (cols * ((n[2] + c - 1) / c) >= cus) {
        *sched = 3;   // loads 1 plane ahead: 1388 us per stage at 512^3 on MI355X, 2 planes 1413, direct 1797
        auto_zc = c;
        break;
      }
  }
  // the column schedule's 32-bit plane offsets
  const long X = f->glob[0], XY = (long)f->glob[0] * f->glob[1];
  if (*sched >= 2 && (XY + 2 * X + 2) * 8 >= (1L << 31)) *sched = 1;
  *zc = g_zchunk.load() > 0 ? g_zchunk.load() : *sched == 1 ? kZChunk : auto_zc;
}

using ColumnFn = void (*)(double*, const double*, const double*, const double*, const double*, const double*,
                          const double*, Tiles, int, int, int, Coefs3, double);

ColumnFn column_fn(int stage, int pf)
{
  static const ColumnFn t[2][3] = {
      {transport_column_kernel<1, kColW, kColNW, kColRY, 2>, transport_column_kernel<2, kColW, kColNW, kColRY, 2>,
       transport_column_kernel<3, kColW, kColNW, kColRY, 2>},
      {transport_column_kernel<1, kColW, kColNW, kColRY, 1>, transport_column_kernel<2, kColW, kColNW, kColRY, 1>,
       transport_column_kernel<3, kColW, kColNW, kColRY, 1>}};
  return t[pf == 2 ? 0 : 1][stage - 1];
}

int launch_stage(const nebo_field_t* out, const nebo_field_t* pn, const nebo_field_t* s,
                 const Operands& o, const nebo_axis_coefs_t& c, double dt, int stage, void* stream)
{
  int sched = 1, zc = kZChunk;
  stage_schedule(out, &sched, &zc);
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  const double* sp = frame_origin(stage == 1 ? pn : s);
  if (sched == 1) {
    const int n[3] = {out->extent[0] - 2, out->extent[1] - 2, out->extent[2] - 2};
    const Box3 B = make_box(n, kTx, kTy, zc);
    const int nb = B.nx * B.ny * B.nz;
    auto fn = stage == 1 ? transport_stage_kernel<1> : stage == 2 ? transport_stage_kernel<2>
                                                                  : transport_stage_kernel<3>;
    hipLaunchKernelGGL(fn, dim3(nb), dim3(kTx * kTy), 0, (hipStream_t)stream, frame_origin(out),
                       frame_origin(pn), sp, frame_origin(o.k), frame_origin(&o.u[0]), frame_origin(&o.u[1]),
                       frame_origin(&o.u[2]), B, X, XY, coefs3(c), dt);
  } else {
    nebo_limits_t L;
    for (int d = 0; d < 3; ++d) {
      L.lo[d] = 0;
      L.hi[d] = out->extent[d] - 2;
    }
    const Tiles T = make_tiles(L, kLanes * kColW, kColRY * kColNW, zc);
    const long nb = (long)T.nx * T.ny * T.nz;
    hipLaunchKernelGGL(column_fn(stage, sched == 2 ? 2 : 1), dim3((unsigned)nb), dim3(kLanes * kColW * kColNW), 0,
                       (hipStream_t)stream, frame_origin(out), frame_origin(pn), sp, frame_origin(o.k),
                       frame_origin(&o.u[0]), frame_origin(&o.u[1]), frame_origin(&o.u[2]), T, zc, X, XY,
                       coefs3(c), dt);
  }
  int rc = check_hip(hipGetLastError(), kCtx);
  if (rc) return rc;
  return fill_ghosts_periodic(out, stream);
}

double stage_weight(int stage) { return stage == 3 ? 2.0 / 3.0 : stage == 2 ? 0.25 : 1.0; }

using AdjColumnFn = void (*)(double*, double*, double*, const double*, const double*, const double*,
                             const double*, const double*, const double*, const double*, Tiles, int, int, int,
                             Coefs3, double, double, double);

AdjColumnFn adjoint_column_fn(int stage, int pf)
{
  static const AdjColumnFn t[2][3] = {
      {transport_adjoint_column_kernel<1, kColW, kColNW, kColRY, 2>,
       transport_adjoint_column_kernel<2, kColW, kColNW, kColRY, 2>,
       transport_adjoint_column_kernel<3, kColW, kColNW, kColRY, 2>},
      {transport_adjoint_column_kernel<1, kColW, kColNW, kColRY, 1>,
       transport_adjoint_column_kernel<2, kColW, kColNW, kColRY, 1>,
       transport_adjoint_column_kernel<3, kColW, kColNW, kColRY, 1>}};
  return t[pf == 2 ? 0 : 1][stage - 1];
}

int launch_adjoint(const nebo_field_t* out, double* acc_out, double* gk, const nebo_field_t* a,
                   const double* acc_in, const nebo_field_t* s, const Operands& o,
                   const nebo_axis_coefs_t& c, double dt, int stage, void* stream)
{
  int sched = 1, zc = kZChunk;
  stage_schedule(out, &sched, &zc);
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  // acc and gk are plain blocks with the field's shape: offset them like the field
  const std::ptrdiff_t off = frame_origin(out) - static_cast<double*>(out->values);
  const double cs = stage_weight(stage);
  const double cdt = cs * dt;
  double* acc_o = acc_out ? acc_out + off : nullptr;
  const double* acc_i = acc_in ? acc_in + off : nullptr;
  if (sched == 1) {
    const int n[3] = {out->extent[0] - 2, out->extent[1] - 2, out->extent[2] - 2};
    const Box3 B = make_box(n, kTx, kTy, zc);
    const int nb = B.nx * B.ny * B.nz;
    auto fn = stage == 1 ? transport_adjoint_kernel<1> : stage == 2 ? transport_adjoint_kernel<2>
                                                                    : transport_adjoint_kernel<3>;
    hipLaunchKernelGGL(fn, dim3(nb), dim3(kTx * kTy), 0, (hipStream_t)stream, frame_origin(out), acc_o, gk + off,
                       frame_origin(a), acc_i, frame_origin(s), frame_origin(o.k), frame_origin(&o.u[0]),
                       frame_origin(&o.u[1]), frame_origin(&o.u[2]), B, X, XY, coefs3(c), dt, cs, cdt);
  } else {
    nebo_limits_t L;
    for (int d = 0; d < 3; ++d) {
      L.lo[d] = 0;
      L.hi[d] = out->extent[d] - 2;
    }
    const Tiles T = make_tiles(L, kLanes * kColW, kColRY * kColNW, zc);
    const long nb = (long)T.nx * T.ny * T.nz;
    hipLaunchKernelGGL(adjoint_column_fn(stage, sched == 2 ? 2 : 1), dim3((unsigned)nb),
                       dim3(kLanes * kColW * kColNW), 0, (hipStream_t)stream, frame_origin(out), acc_o, gk + off,
                       frame_origin(a), acc_i, frame_origin(s), frame_origin(o.k), frame_origin(&o.u[0]),
                       frame_origin(&o.u[1]), frame_origin(&o.u[2]), T, zc, X, XY, coefs3(c), dt, cs, cdt);
  }
  int rc = check_hip(hipGetLastError(), kCtx);
  if (rc) return rc;
  return fill_ghosts_periodic(out, stream);
}

// k, u[0..2] and every state field share ref's periodic window.
int check_operands(const nebo_field_t* ref, const nebo_field_t* k, const nebo_field_t* u)
{
  int rc;
  if ((rc = check_periodic_operand(ref, ref, kCtx)) || (rc = check_periodic_operand(k, ref, kCtx)))
    return rc;
  for (int d = 0; d < 3; ++d)
    if ((rc = check_periodic_operand(&u[d], ref, kCtx))) return rc;
  return NEBO_OK;
}

// One RK3 step phi_n -> phi_next through s1, s2 (all periodic-filled).
int rk3_step(double* next, double* s1, double* s2, const double* phin, const nebo_field_t* ref,
             const Operands& o, const nebo_axis_coefs_t& c, double dt, void* stream)
{
  const nebo_field_t P = with_values(ref, const_cast<double*>(phin));
  const nebo_field_t S1 = with_values(ref, s1), S2 = with_values(ref, s2), N = with_values(ref, next);
  int rc = launch_stage(&S1, &P, &P, o, c, dt, 1, stream);
  if (!rc) rc = launch_stage(&S2, &P, &S1, o, c, dt, 2, stream);
  if (!rc) rc = launch_stage(&N, &P, &S2, o, c, dt, 3, stream);
  return rc;
}

}  // namespace
}  // namespace nebo

using namespace nebo;

extern "C" int nebo_transport_tune(int32_t schedule, int32_t zchunk)
{
  clear_error();
  if (schedule < 0 || schedule > 3) return set_error(NEBO_ERR_INVALID, kCtx, "schedule: 0 to 3");
  if (zchunk < 0 || zchunk > 4096)
    return set_error(NEBO_ERR_INVALID, kCtx, "z chunk: 0 (the schedule's default) or 1 to 4096 planes");
  g_schedule.store(schedule);
  g_zchunk.store(zchunk);
  return NEBO_OK;
}

extern "C" int nebo_transport_schedule(const nebo_field_t* field, int32_t* schedule, int32_t* zchunk)
{
  clear_error();
  if (!field) return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  int sched = 1, zc = kZChunk;
  stage_schedule(field, &sched, &zc);
  if (schedule) *schedule = sched;
  if (zchunk) *zchunk = zc;
  return NEBO_OK;
}

extern "C" int nebo_transport_stage(const nebo_field_t* out, const nebo_field_t* phin,
                                    const nebo_field_t* s, const nebo_field_t* k,
                                    const nebo_field_t* u, const nebo_axis_coefs_t* c, double dt,
                                    int stage, void* stream)
{
  clear_error();
  if (!out || !phin || !k || !u || !c || (stage != 1 && !s))
    return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  if (stage < 1 || stage > 3) return set_error(NEBO_ERR_INVALID, kCtx, "stage must be 1, 2 or 3");
  int rc;
  if ((rc = check_operands(out, k, u)) || (rc = check_periodic_operand(phin, out, kCtx))) return rc;
  if (stage != 1 && (rc = check_periodic_operand(s, out, kCtx))) return rc;
  if (out->values == phin->values || out->values == k->values || (stage != 1 && out->values == s->values))
    return set_error(NEBO_ERR_INVALID, kCtx, "output must not alias an input");
  for (int d = 0; d < 3; ++d)
    if (out->values == u[d].values) return set_error(NEBO_ERR_INVALID, kCtx, "output must not alias an input");
  return launch_stage(out, phin, s, Operands{u, k}, *c, dt, stage, stream);
}

extern "C" int nebo_transport_adjoint_stage(const nebo_field_t* out, double* acc_out, double* gk,
                                            const nebo_field_t* a, const double* acc_in,
                                            const nebo_field_t* s, const nebo_field_t* k,
                                            const nebo_field_t* u, const nebo_axis_coefs_t* c,
                                            double dt, int stage, void* stream)
{
  clear_error();
  if (!out || !gk || !a || !s || !k || !u || !c) return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  if (stage < 1 || stage > 3) return set_error(NEBO_ERR_INVALID, kCtx, "stage must be 1, 2 or 3");
  if ((stage != 1 && !acc_out) || (stage != 3 && !acc_in))
    return set_error(NEBO_ERR_INVALID, kCtx, "stages 2 and 3 write acc_out; stages 1 and 2 read acc_in");
  int rc;
  if ((rc = check_operands(out, k, u)) || (rc = check_periodic_operand(a, out, kCtx)) ||
      (rc = check_periodic_operand(s, out, kCtx)))
    return rc;
  const void* outs[3] = {out->values, acc_out, gk};
  const void* ins[6] = {a->values, acc_in, s->values, k->values, u[0].values, u[1].values};
  for (const void* w : outs)
    for (const void* r : ins)
      if (w && r && w == r) return set_error(NEBO_ERR_INVALID, kCtx, "outputs must not alias inputs");
  if (out->values == u[2].values || (acc_out && acc_out == gk) || (void*)gk == out->values)
    return set_error(NEBO_ERR_INVALID, kCtx, "outputs must not alias inputs or each other");
  return launch_adjoint(out, stage == 1 ? nullptr : acc_out, gk, a, stage == 3 ? nullptr : acc_in, s,
                        Operands{u, k}, *c, dt, stage, stream);
}

extern "C" int nebo_transport_forward(const nebo_field_t* phiT, const nebo_field_t* phi0,
                                      const nebo_field_t* k, const nebo_field_t* u,
                                      const nebo_axis_coefs_t* c, double dt, int32_t n_steps,
                                      int32_t every, double* ckpt, void* stream)
{
  clear_error();
  if (!phiT || !phi0 || !k || !u || !c || !ckpt) return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  if (n_steps < 1 || every < 1 || n_steps % every)
    return set_error(NEBO_ERR_INVALID, kCtx, "n_steps must be a positive multiple of the checkpoint interval");
  int rc;
  if ((rc = check_operands(phi0, k, u)) || (rc = check_periodic_operand(phiT, phi0, kCtx))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const size_t np = field_npts(phi0), bytes = np * sizeof(double);
  double* tmp = nullptr;   // s1, s2 and two ping-pong states
  if ((rc = check_hip(hipMallocAsync((void**)&tmp, 4 * bytes, st), kCtx))) return rc;
  const Operands o{u, k};
  nebo_field_t ck0 = with_values(phi0, ckpt);
  rc = check_hip(hipMemcpyAsync(ckpt, phi0->values, bytes, hipMemcpyDeviceToDevice, st), kCtx);
  if (!rc) rc = fill_ghosts_periodic(&ck0, stream);
  double* cur = ckpt;
  for (int n = 0; n < n_steps && !rc; ++n) {
    double* dst;
    if ((n + 1) % every == 0) dst = ckpt + (size_t)((n + 1) / every) * np;
    else dst = (cur == tmp + 2 * np) ? tmp + 3 * np : tmp + 2 * np;
    rc = rk3_step(dst, tmp, tmp + np, cur, phi0, o, *c, dt, stream);
    cur = dst;
  }
  if (!rc && cur != phiT->values)
    rc = check_hip(hipMemcpyAsync(phiT->values, cur, bytes, hipMemcpyDeviceToDevice, st), kCtx);
  (void)hipFreeAsync(tmp, st);
  return rc;
}

extern "C" int nebo_transport_backward(const nebo_field_t* lam0, double* gk, const nebo_field_t* lamT,
                                       const nebo_field_t* k, const nebo_field_t* u,
                                       const nebo_axis_coefs_t* c, double dt, int32_t n_steps,
                                       int32_t every, const double* ckpt, double* work, void* stream)
{
  clear_error();
  if (!lam0 || !gk || !lamT || !k || !u || !c || !ckpt || !work)
    return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  if (n_steps < 1 || every < 1 || n_steps % every)
    return set_error(NEBO_ERR_INVALID, kCtx, "n_steps must be a positive multiple of the checkpoint interval");
  int rc;
  if ((rc = check_operands(lamT, k, u)) || (rc = check_periodic_operand(lam0, lamT, kCtx))) return rc;
  if (lam0->values == lamT->values) return set_error(NEBO_ERR_INVALID, kCtx, "lam0 must not alias lamT");
  hipStream_t st = (hipStream_t)stream;
  const size_t np = field_npts(lamT), bytes = np * sizeof(double);
  const Operands o{u, k};
  // work: states (phi_n, s1, s2) of one segment's steps, then lam, mu2, mu1, acc, acc2
  double* states = work;
  double* lamA = work + (size_t)3 * every * np;
  double* mu2 = lamA + np;
  double* mu1 = mu2 + np;
  double* acc = mu1 + np;
  double* acc2 = acc + np;
  double* lbuf[2] = {lamA, static_cast<double*>(lam0->values)};
  int cur = 0;
  nebo_field_t L0 = with_values(lamT, lbuf[0]);
  rc = check_hip(hipMemcpyAsync(lbuf[0], lamT->values, bytes, hipMemcpyDeviceToDevice, st), kCtx);
  if (!rc) rc = fill_ghosts_periodic(&L0, stream);
  if (!rc) rc = check_hip(hipMemsetAsync(gk, 0, field_npts(k) * sizeof(double), st), kCtx);
  const int nseg = n_steps / every;
  for (int seg = nseg - 1; seg >= 0 && !rc; --seg) {
    // recompute the segment: states[t] = (phi_t, s1_t, s2_t), phi_0 = the checkpoint
    rc = check_hip(hipMemcpyAsync(states, ckpt + (size_t)seg * np, bytes, hipMemcpyDeviceToDevice, st), kCtx);
    for (int t = 0; t < every && !rc; ++t) {
      double* phit = states + (size_t)3 * t * np;
      double* nxt = t + 1 < every ? states + (size_t)3 * (t + 1) * np : mu2;   // last next is not kept
      rc = rk3_step(nxt, phit + np, phit + 2 * np, phit, lamT, o, *c, dt, stream);
    }
    for (int t = every - 1; t >= 0 && !rc; --t) {
      const double* phit = states + (size_t)3 * t * np;
      const nebo_field_t Lin = with_values(lamT, lbuf[cur]), Lout = with_values(lamT, lbuf[1 - cur]);
      const nebo_field_t M2 = with_values(lamT, mu2), M1 = with_values(lamT, mu1);
      const nebo_field_t Pn = with_values(lamT, const_cast<double*>(phit));
      const nebo_field_t S1 = with_values(lamT, const_cast<double*>(phit + np));
      const nebo_field_t S2 = with_values(lamT, const_cast<double*>(phit + 2 * np));
      rc = launch_adjoint(&M2, acc, gk, &Lin, nullptr, &S2, o, *c, dt, 3, stream);
      if (!rc) rc = launch_adjoint(&M1, acc2, gk, &M2, acc, &S1, o, *c, dt, 2, stream);
      if (!rc) rc = launch_adjoint(&Lout, nullptr, gk, &M1, acc2, &Pn, o, *c, dt, 1, stream);
      cur = 1 - cur;
    }
  }
  if (!rc && lbuf[cur] != lam0->values)
    rc = check_hip(hipMemcpyAsync(lam0->values, lbuf[cur], bytes, hipMemcpyDeviceToDevice, st), kCtx);
  return rc;
}
