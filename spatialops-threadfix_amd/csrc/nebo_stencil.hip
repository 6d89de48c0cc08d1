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
// Work split: a workgroup of 64*TY threads covers a 64 (x) by TY (y) column tile
// and marches through a chunk of ZC z planes, keeping the column's z neighbours
// and the shared z face in registers; x and y neighbours come from the L1/L2
// caches.  Halo rows and planes are re-read by the neighbouring tile, so larger
// TY and ZC cut the traffic beyond L2 ((TY+2)/TY * (ZC+2)/ZC for the operands)
// at the cost of fewer workgroups; the best shape is measured on the device
// (nebo_stencil_autotune).  Workgroups are numbered so that consecutive tiles,
// which share halo rows, land on the same XCD and therefore the same L2.

#include "nebo_internal.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <sstream>

namespace nebo {
namespace {

constexpr int kTx = 64;

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

Tiles make_tiles(const nebo_limits_t& L, int ty, int zc)
{
  Tiles t;
  for (int d = 0; d < 3; ++d) { t.lo[d] = L.lo[d]; t.hi[d] = L.hi[d]; }
  t.nx = (L.hi[0] - L.lo[0] + kTx - 1) / kTx;
  t.ny = (L.hi[1] - L.lo[1] + ty - 1) / ty;
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

// FORM 0: (Dx + Dy) + Dz;  FORM 1: ((-Dx) - Dy) - Dz;  FORM 2: phi + dt*((Dx + Dy) + Dz)
template <int FORM, int TY, int ZC>
__global__ __launch_bounds__(kTx * TY) void divkgrad_kernel(
    double* __restrict__ r, const double* __restrict__ k, const double* __restrict__ p,
    Tiles T, int X, int XY, nebo_axis_coefs_t c, double dt)
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
template <int TY, int ZC>
__global__ __launch_bounds__(kTx * TY) void lapl_euler_kernel(
    double* __restrict__ out, const double* __restrict__ p, Tiles T, int X, int XY,
    nebo_axis_coefs_t c, double s)
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

// ---- tile table -------------------------------------------------------------

struct TileShape { int ty, zc; };
constexpr TileShape kShapes[] = {{4, 16}, {4, 32}, {4, 64}, {8, 16}, {8, 32}, {8, 64},
                                 {16, 16}, {16, 32}, {16, 64}, {2, 64}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);
constexpr int kDefaultShape = 0;

using DivkFn = void (*)(double*, const double*, const double*, Tiles, int, int, nebo_axis_coefs_t, double);
using LaplFn = void (*)(double*, const double*, Tiles, int, int, nebo_axis_coefs_t, double);

template <int FORM, int S>
constexpr DivkFn divk_fn() { return divkgrad_kernel<FORM, kShapes[S].ty, kShapes[S].zc>; }
template <int S>
constexpr LaplFn lapl_fn() { return lapl_euler_kernel<kShapes[S].ty, kShapes[S].zc>; }

template <int FORM>
DivkFn divk_table(int s)
{
  static const DivkFn t[kNumShapes] = {divk_fn<FORM, 0>(), divk_fn<FORM, 1>(), divk_fn<FORM, 2>(),
                                       divk_fn<FORM, 3>(), divk_fn<FORM, 4>(), divk_fn<FORM, 5>(),
                                       divk_fn<FORM, 6>(), divk_fn<FORM, 7>(), divk_fn<FORM, 8>(),
                                       divk_fn<FORM, 9>()};
  return t[s];
}

LaplFn lapl_table(int s)
{
  static const LaplFn t[kNumShapes] = {lapl_fn<0>(), lapl_fn<1>(), lapl_fn<2>(), lapl_fn<3>(),
                                       lapl_fn<4>(), lapl_fn<5>(), lapl_fn<6>(), lapl_fn<7>(),
                                       lapl_fn<8>(), lapl_fn<9>()};
  return t[s];
}

static_assert(kNumShapes == 10, "tables list every shape");

std::atomic<int> g_divk_shape{kDefaultShape};
std::atomic<int> g_lapl_shape{kDefaultShape};

int find_shape(int ty, int zc)
{
  for (int s = 0; s < kNumShapes; ++s)
    if (kShapes[s].ty == ty && kShapes[s].zc == zc) return s;
  return -1;
}

int launch_divk_shape(int shape, int form, const nebo_field_t* r, const nebo_field_t* k,
                      const nebo_field_t* phi, const nebo_axis_coefs_t& c, double dt,
                      const nebo_limits_t& lim, void* stream)
{
  const TileShape sh = kShapes[shape];
  Tiles T = make_tiles(lim, sh.ty, sh.zc);
  const int nb = T.nx * T.ny * T.nz;
  const int X = r->glob[0], XY = r->glob[0] * r->glob[1];
  DivkFn fn = form == 0 ? divk_table<0>(shape) : form == 1 ? divk_table<1>(shape) : divk_table<2>(shape);
  hipLaunchKernelGGL(fn, dim3(nb), dim3(kTx * sh.ty), 0, (hipStream_t)stream, frame_origin(r),
                     frame_origin(k), frame_origin(phi), T, X, XY, c, dt);
  return check_hip(hipGetLastError(), "Nebo Assignment");
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
  return launch_divk_shape(g_divk_shape.load(), form, r, k, phi, c, 0.0, lim, stream);
}

int launch_diffusion_euler(const nebo_field_t* out, const nebo_field_t* k, const nebo_field_t* phi,
                           const nebo_axis_coefs_t& c, double dt, const nebo_limits_t& lim,
                           void* stream)
{
  return launch_divk_shape(g_divk_shape.load(), 2, out, k, phi, c, dt, lim, stream);
}

int launch_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                      const nebo_axis_coefs_t& c, double sc, const nebo_limits_t& lim,
                      void* stream)
{
  const TileShape sh = kShapes[g_lapl_shape.load()];
  Tiles T = make_tiles(lim, sh.ty, sh.zc);
  const int nb = T.nx * T.ny * T.nz;
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  hipLaunchKernelGGL(lapl_table(g_lapl_shape.load()), dim3(nb), dim3(kTx * sh.ty), 0, (hipStream_t)stream,
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

extern "C" int nebo_stencil_set_tile(int kernel, int32_t ty, int32_t zchunk)
{
  clear_error();
  const int s = find_shape(ty, zchunk);
  if (s < 0 || (kernel != 0 && kernel != 1)) {
    std::ostringstream m;
    m << "no stencil kernel instantiated for kernel " << kernel << ", tile 64x" << ty << "x" << zchunk;
    return set_error(NEBO_ERR_INVALID, "Nebo Stencil Tile", m.str());
  }
  (kernel == 0 ? g_divk_shape : g_lapl_shape).store(s);
  return NEBO_OK;
}

extern "C" int nebo_stencil_get_tile(int kernel, int32_t* ty, int32_t* zchunk)
{
  clear_error();
  if (kernel != 0 && kernel != 1) return set_error(NEBO_ERR_INVALID, "Nebo Stencil Tile", "unknown kernel");
  const TileShape sh = kShapes[(kernel == 0 ? g_divk_shape : g_lapl_shape).load()];
  if (ty) *ty = sh.ty;
  if (zchunk) *zchunk = sh.zc;
  return NEBO_OK;
}

extern "C" int nebo_stencil_autotune(const nebo_field_t* r, const nebo_field_t* k,
                                     const nebo_field_t* phi, const nebo_axis_coefs_t* c,
                                     void* stream, int32_t* ty, int32_t* zchunk, float* times_ms)
{
  static const char* ctx = "Nebo Stencil Tile";
  clear_error();
  if (!r || !k || !phi || !c) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  int rc;
  if ((rc = check_same_shape(r, k, ctx)) || (rc = check_same_shape(r, phi, ctx))) return rc;
  if (r->values == k->values || r->values == phi->values)
    return set_error(NEBO_ERR_INVALID, ctx, "result must not alias an operand");
  nebo_limits_t L;
  int empty = 0;
  stencil_range(k, &L, &empty);
  if (empty) return set_error(NEBO_ERR_INVALID, ctx, "no interior cells to time");
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  if ((rc = check_hip(hipEventCreate(&e0), ctx))) return rc;
  if ((rc = check_hip(hipEventCreate(&e1), ctx))) return rc;
  int best = g_divk_shape.load();
  float best_ms = 1e30f;
  for (int sh = 0; sh < kNumShapes && !rc; ++sh) {
    rc = launch_divk_shape(sh, 0, r, k, phi, *c, 0.0, L, stream);   // warm
    if (!rc) rc = check_hip(hipEventRecord(e0, s), ctx);
    for (int rep = 0; rep < 5 && !rc; ++rep) rc = launch_divk_shape(sh, 0, r, k, phi, *c, 0.0, L, stream);
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
  g_divk_shape.store(best);
  if (ty) *ty = kShapes[best].ty;
  if (zchunk) *zchunk = kShapes[best].zc;
  return NEBO_OK;
}

extern "C" int nebo_stencil_tiles(int32_t* n, int32_t* ty, int32_t* zchunk)
{
  clear_error();
  if (n) *n = kNumShapes;
  for (int s = 0; s < kNumShapes; ++s) {
    if (ty) ty[s] = kShapes[s].ty;
    if (zchunk) zchunk[s] = kShapes[s].zc;
  }
  return NEBO_OK;
}
