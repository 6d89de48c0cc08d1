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
// Work split: a workgroup of 256 threads covers a 64 (x) by 4 (y) column tile
// and marches through a chunk of z planes, keeping the column's z neighbours and
// the shared z face in registers.  x and y neighbours come from the L1/L2
// caches.  Workgroups are numbered so that consecutive tiles (which share halo
// rows) land on the same XCD and therefore the same L2.

#include "nebo_internal.h"

#include <hip/hip_runtime.h>

namespace nebo {
namespace {

constexpr int kTx = 64;
constexpr int kTy = 4;
constexpr int kZChunk = 16;

// Hardware dispatches workgroup b to XCD b % 8.  Map b to a logical tile so that
// each XCD receives one contiguous run of logical tiles.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, k = b >> 3;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

struct Tiles {
  int lo[3], hi[3];
  int nx, ny, nz;  // tile counts
};

Tiles make_tiles(const nebo_limits_t& L, int zc)
{
  Tiles t;
  for (int d = 0; d < 3; ++d) { t.lo[d] = L.lo[d]; t.hi[d] = L.hi[d]; }
  t.nx = (L.hi[0] - L.lo[0] + kTx - 1) / kTx;
  t.ny = (L.hi[1] - L.lo[1] + kTy - 1) / kTy;
  t.nz = (L.hi[2] - L.lo[2] + zc - 1) / zc;
  return t;
}

// interp(k)*grad(phi) on one face: (k_lo*i0 + k_hi*i1) * (p_lo*g0 + p_hi*g1)
__device__ __forceinline__ double face(double klo, double khi, double plo, double phi,
                                       double i0, double i1, double g0, double g1)
{
  const double ki = klo * i0 + khi * i1;
  const double gr = plo * g0 + phi * g1;
  return ki * gr;
}

template <int FORM>
__global__ __launch_bounds__(kTx * kTy) void divkgrad_kernel(
    double* __restrict__ r, const double* __restrict__ k, const double* __restrict__ p,
    Tiles T, int X, int XY, nebo_axis_coefs_t c, double dt)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int x = T.lo[0] + tx * kTx + (threadIdx.x & (kTx - 1));
  const int y = T.lo[1] + ty * kTy + (threadIdx.x / kTx);
  if (x >= T.hi[0] || y >= T.hi[1]) return;
  const int z0 = T.lo[2] + tz * kZChunk;
  const int z1 = min(z0 + kZChunk, T.hi[2]);

  const double ix0 = c.interp[0][0], ix1 = c.interp[0][1], gx0 = c.grad[0][0], gx1 = c.grad[0][1];
  const double iy0 = c.interp[1][0], iy1 = c.interp[1][1], gy0 = c.grad[1][0], gy1 = c.grad[1][1];
  const double iz0 = c.interp[2][0], iz1 = c.interp[2][1], gz0 = c.grad[2][0], gz1 = c.grad[2][1];
  const double dx0 = c.div[0][0], dx1 = c.div[0][1];
  const double dy0 = c.div[1][0], dy1 = c.div[1][1];
  const double dz0 = c.div[2][0], dz1 = c.div[2][1];

  int i = x + X * y + XY * z0;
  double pc = p[i], kc = k[i];
  double fz_lo = face(k[i - XY], kc, p[i - XY], pc, iz0, iz1, gz0, gz1);
  for (int z = z0; z < z1; ++z, i += XY) {
    const double pz = p[i + XY], kz = k[i + XY];
    const double pxm = p[i - 1], pxp = p[i + 1], kxm = k[i - 1], kxp = k[i + 1];
    const double pym = p[i - X], pyp = p[i + X], kym = k[i - X], kyp = k[i + X];
    const double fx_lo = face(kxm, kc, pxm, pc, ix0, ix1, gx0, gx1);
    const double fx_hi = face(kc, kxp, pc, pxp, ix0, ix1, gx0, gx1);
    const double fy_lo = face(kym, kc, pym, pc, iy0, iy1, gy0, gy1);
    const double fy_hi = face(kc, kyp, pc, pyp, iy0, iy1, gy0, gy1);
    const double fz_hi = face(kc, kz, pc, pz, iz0, iz1, gz0, gz1);
    const double Dx = fx_lo * dx0 + fx_hi * dx1;
    const double Dy = fy_lo * dy0 + fy_hi * dy1;
    const double Dz = fz_lo * dz0 + fz_hi * dz1;
    double v;
    if constexpr (FORM == 0) v = (Dx + Dy) + Dz;
    else if constexpr (FORM == 1) v = ((-Dx) - Dy) - Dz;
    else v = pc + dt * ((Dx + Dy) + Dz);   // explicit Euler step of the diffusion equation
    __builtin_nontemporal_store(v, &r[i]);
    pc = pz;
    kc = kz;
    fz_lo = fz_hi;
  }
}

// out = phi + s*((Lx + Ly) + Lz),  L_d = (phi[-1]*g0 + phi[0]*g1)*d0 + (phi[0]*g0 + phi[+1]*g1)*d1
__global__ __launch_bounds__(kTx * kTy) void lapl_euler_kernel(
    double* __restrict__ out, const double* __restrict__ p, Tiles T, int X, int XY,
    nebo_axis_coefs_t c, double s)
{
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % T.nx;
  const int rest = t / T.nx;
  const int ty = rest % T.ny;
  const int tz = rest / T.ny;
  const int x = T.lo[0] + tx * kTx + (threadIdx.x & (kTx - 1));
  const int y = T.lo[1] + ty * kTy + (threadIdx.x / kTx);
  if (x >= T.hi[0] || y >= T.hi[1]) return;
  const int z0 = T.lo[2] + tz * kZChunk;
  const int z1 = min(z0 + kZChunk, T.hi[2]);
  const double gx0 = c.grad[0][0], gx1 = c.grad[0][1], gy0 = c.grad[1][0], gy1 = c.grad[1][1];
  const double gz0 = c.grad[2][0], gz1 = c.grad[2][1];
  const double dx0 = c.div[0][0], dx1 = c.div[0][1], dy0 = c.div[1][0], dy1 = c.div[1][1];
  const double dz0 = c.div[2][0], dz1 = c.div[2][1];

  int i = x + X * y + XY * z0;
  double pc = p[i];
  double gz_lo = p[i - XY] * gz0 + pc * gz1;
  for (int z = z0; z < z1; ++z, i += XY) {
    const double pz = p[i + XY];
    const double pxm = p[i - 1], pxp = p[i + 1], pym = p[i - X], pyp = p[i + X];
    const double gz_hi = pc * gz0 + pz * gz1;
    const double Lx = (pxm * gx0 + pc * gx1) * dx0 + (pc * gx0 + pxp * gx1) * dx1;
    const double Ly = (pym * gy0 + pc * gy1) * dy0 + (pc * gy0 + pyp * gy1) * dy1;
    const double Lz = gz_lo * dz0 + gz_hi * dz1;
    const double v = pc + s * ((Lx + Ly) + Lz);
    __builtin_nontemporal_store(v, &out[i]);
    pc = pz;
    gz_lo = gz_hi;
  }
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
  Tiles T = make_tiles(lim, kZChunk);
  const int nb = T.nx * T.ny * T.nz;
  const int X = r->glob[0], XY = r->glob[0] * r->glob[1];
  hipStream_t s = (hipStream_t)stream;
  if (form == 0)
    hipLaunchKernelGGL(divkgrad_kernel<0>, dim3(nb), dim3(kTx * kTy), 0, s, frame_origin(r),
                       frame_origin(k), frame_origin(phi), T, X, XY, c, 0.0);
  else
    hipLaunchKernelGGL(divkgrad_kernel<1>, dim3(nb), dim3(kTx * kTy), 0, s, frame_origin(r),
                       frame_origin(k), frame_origin(phi), T, X, XY, c, 0.0);
  return check_hip(hipGetLastError(), "Nebo Assignment");
}

int launch_diffusion_euler(const nebo_field_t* out, const nebo_field_t* k, const nebo_field_t* phi,
                           const nebo_axis_coefs_t& c, double dt, const nebo_limits_t& lim,
                           void* stream)
{
  Tiles T = make_tiles(lim, kZChunk);
  const int nb = T.nx * T.ny * T.nz;
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  hipLaunchKernelGGL(divkgrad_kernel<2>, dim3(nb), dim3(kTx * kTy), 0, (hipStream_t)stream,
                     frame_origin(out), frame_origin(k), frame_origin(phi), T, X, XY, c, dt);
  return check_hip(hipGetLastError(), "Nebo Diffusion");
}

int launch_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                      const nebo_axis_coefs_t& c, double sc, const nebo_limits_t& lim,
                      void* stream)
{
  Tiles T = make_tiles(lim, kZChunk);
  const int nb = T.nx * T.ny * T.nz;
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  hipLaunchKernelGGL(lapl_euler_kernel, dim3(nb), dim3(kTx * kTy), 0, (hipStream_t)stream,
                     frame_origin(out), frame_origin(phi), T, X, XY, c, sc);
  return check_hip(hipGetLastError(), "Nebo Assignment");
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

int interior(const nebo_field_t* f, nebo_limits_t* L, const char* ctx)
{
  for (int d = 0; d < 3; ++d) {
    if (f->valid_minus[d] < 1 || f->valid_plus[d] < 1)
      return set_error(NEBO_ERR_GHOST, ctx, "the stencil needs one valid ghost layer on each face");
    L->lo[d] = 0;
    L->hi[d] = f->extent[d] - f->valid_minus[d] - f->valid_plus[d];
  }
  return NEBO_OK;
}

}  // namespace
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
  if ((rc = interior(k, &L, ctx))) return rc;
  for (int d = 0; d < 3; ++d) L.hi[d] = r->extent[d] - r->valid_minus[d] - r->valid_plus[d];
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
  if ((rc = interior(phi, &L, ctx))) return rc;
  return launch_lapl_euler(out, phi, *c, s, L, stream);
}
