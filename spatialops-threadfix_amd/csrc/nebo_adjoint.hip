// nebo_adjoint.hip -- adjoint of the periodic var-coefficient diffusion step and
// the checkpointed forward / reverse sweeps.
//
// Forward step (interior, periodic ghosts):  phi' = phi + dt * D(I(k) .* G phi)
// with G = Gradient (face i = (phi_{i-1}*g0 + phi_i*g1)), I = Interpolant
// (k_{i-1}*i0 + k_i*i1) and D = Divergence (cell c = F_c*d0 + F_{c+1}*d1).
// Its transpose, for lam = dL/dphi':
//   w_i   = (D^T lam)_i   = lam_i*d0 + lam_{i-1}*d1                (faces)
//   a_c   = (G^T (kf.*w))_c = (kf*w)_{c+1}*g0 + (kf*w)_c*g1        (cells)
//   b_c   = (I^T (w.*Gphi))_c = (w*gp)_{c+1}*i0 + (w*gp)_c*i1      (cells)
//   dL/dphi = lam + dt*((ax + ay) + az),  dL/dk += dt*((bx + by) + bz)
// Periodicity enters through the ghost layers of lam, phi and k, which hold the
// wrapped interior values, so the faces at both ends of an axis are the same
// physical face seen from its two cells.
//
// One fused pass per adjoint step: reads lam, phi, k and gk, writes lam' and gk
// (48 bytes per cell); the z faces are carried in registers.

#include "nebo_internal.h"
#include "nebo_device.h"

#include <hip/hip_runtime.h>

#include <sstream>

namespace nebo {

int check_periodic_operand(const nebo_field_t* f, const nebo_field_t* ref, const char* ctx)
{
  for (int d = 0; d < 3; ++d) {
    if (f->glob[d] != ref->glob[d] || f->extent[d] != ref->extent[d] || f->offset[d] != ref->offset[d])
      return set_error(NEBO_ERR_INVALID, ctx, "operands must share one memory window");
    if (f->ghost_minus[d] != 1 || f->ghost_plus[d] != 1 || f->valid_minus[d] < 1 || f->valid_plus[d] < 1)
      return set_error(NEBO_ERR_GHOST, ctx, "periodic operands need exactly one ghost layer per face");
    if (f->extent[d] - 2 < 1)
      return set_error(NEBO_ERR_INVALID, ctx, "empty interior");
  }
  if ((long)f->glob[0] * f->glob[1] * f->glob[2] >= (1L << 31))
    return set_error(NEBO_ERR_TOO_LARGE, ctx, "field has 2^31 or more points");
  if (f->device < 0 || f->device != ref->device)
    return set_error(NEBO_ERR_DEVICE, ctx, "operands must be resident on one GPU");
  return NEBO_OK;
}

nebo_limits_t interior_of(const nebo_field_t* f)
{
  nebo_limits_t L;
  for (int d = 0; d < 3; ++d) {
    L.lo[d] = 0;
    L.hi[d] = f->extent[d] - 2;
  }
  return L;
}

nebo_field_t with_values(const nebo_field_t* f, void* v)
{
  nebo_field_t g = *f;
  g.values = v;
  return g;
}

size_t field_npts(const nebo_field_t* f) { return (size_t)f->glob[0] * f->glob[1] * f->glob[2]; }

namespace {

constexpr int kTx = 64;
constexpr int kTy = 4;
constexpr int kZChunk = 16;

struct FaceQ {
  double kw;  // kf * w
  double wg;  // w * gp
};

// face between cells lo and hi along one axis
__device__ __forceinline__ FaceQ face_q(double lam_lo, double lam_hi, double p_lo, double p_hi,
                                        double k_lo, double k_hi, double d0, double d1, double g0,
                                        double g1, double i0, double i1)
{
  const double w = lam_hi * d0 + lam_lo * d1;
  const double gp = p_lo * g0 + p_hi * g1;
  const double kf = k_lo * i0 + k_hi * i1;
  return FaceQ{kf * w, w * gp};
}

__global__ __launch_bounds__(kTx * kTy) void adjoint_kernel(
    double* __restrict__ lam_out, double* __restrict__ gk, const double* __restrict__ lam,
    const double* __restrict__ p, const double* __restrict__ k, Box3 B, int X, int XY,
    nebo_axis_coefs_t c, double dt)
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
  const double i0x = c.interp[0][0], i1x = c.interp[0][1], g0x = c.grad[0][0], g1x = c.grad[0][1];
  const double i0y = c.interp[1][0], i1y = c.interp[1][1], g0y = c.grad[1][0], g1y = c.grad[1][1];
  const double i0z = c.interp[2][0], i1z = c.interp[2][1], g0z = c.grad[2][0], g1z = c.grad[2][1];
  const double d0x = c.div[0][0], d1x = c.div[0][1], d0y = c.div[1][0], d1y = c.div[1][1];
  const double d0z = c.div[2][0], d1z = c.div[2][1];

  int i = x + X * y + XY * z0;
  double lc = lam[i], pc = p[i], kc = k[i];
  FaceQ zlo = face_q(lam[i - XY], lc, p[i - XY], pc, k[i - XY], kc, d0z, d1z, g0z, g1z, i0z, i1z);
  for (int z = z0; z < z1; ++z, i += XY) {
    const double lz = lam[i + XY], pz = p[i + XY], kz = k[i + XY];
    const FaceQ xlo = face_q(lam[i - 1], lc, p[i - 1], pc, k[i - 1], kc, d0x, d1x, g0x, g1x, i0x, i1x);
    const FaceQ xhi = face_q(lc, lam[i + 1], pc, p[i + 1], kc, k[i + 1], d0x, d1x, g0x, g1x, i0x, i1x);
    const FaceQ ylo = face_q(lam[i - X], lc, p[i - X], pc, k[i - X], kc, d0y, d1y, g0y, g1y, i0y, i1y);
    const FaceQ yhi = face_q(lc, lam[i + X], pc, p[i + X], kc, k[i + X], d0y, d1y, g0y, g1y, i0y, i1y);
    const FaceQ zhi = face_q(lc, lz, pc, pz, kc, kz, d0z, d1z, g0z, g1z, i0z, i1z);
    const double ax = xhi.kw * g0x + xlo.kw * g1x;
    const double ay = yhi.kw * g0y + ylo.kw * g1y;
    const double az = zhi.kw * g0z + zlo.kw * g1z;
    const double bx = xhi.wg * i0x + xlo.wg * i1x;
    const double by = yhi.wg * i0y + ylo.wg * i1y;
    const double bz = zhi.wg * i0z + zlo.wg * i1z;
    lam_out[i] = lc + dt * ((ax + ay) + az);
    gk[i] = gk[i] + dt * ((bx + by) + bz);
    lc = lz;
    pc = pz;
    kc = kz;
    zlo = zhi;
  }
}

const char* kCtx = "Nebo Adjoint";

int check_periodic_field(const nebo_field_t* f, const nebo_field_t* ref)
{
  return check_periodic_operand(f, ref, kCtx);
}

// out = phi + dt*div(interp(k)*grad(phi)), then out's periodic ghosts.  phi's
// ghosts must already hold the periodic images.
int step_filled(const nebo_field_t* out, const nebo_field_t* phi, const nebo_field_t* k,
                const nebo_axis_coefs_t& c, double dt, void* stream)
{
  int rc = launch_diffusion_euler(out, k, phi, c, dt, interior_of(out), stream);
  if (rc) return rc;
  return fill_ghosts_periodic(out, stream);
}

int adjoint_filled(const nebo_field_t* lam_out, double* gk, const nebo_field_t* lam,
                   const nebo_field_t* phi, const nebo_field_t* k, const nebo_axis_coefs_t& c,
                   double dt, void* stream)
{
  Box3 B;
  for (int d = 0; d < 3; ++d) B.n[d] = lam_out->extent[d] - 2;
  B.nx = (B.n[0] + kTx - 1) / kTx;
  B.ny = (B.n[1] + kTy - 1) / kTy;
  B.nz = (B.n[2] + kZChunk - 1) / kZChunk;
  B.zc = kZChunk;
  const int nb = B.nx * B.ny * B.nz;
  const int X = lam_out->glob[0], XY = lam_out->glob[0] * lam_out->glob[1];
  // gk has k's shape: the same frame offset as k
  double* gko = gk + (frame_origin(k) - static_cast<double*>(k->values));
  hipLaunchKernelGGL(adjoint_kernel, dim3(nb), dim3(kTx * kTy), 0, (hipStream_t)stream,
                     frame_origin(lam_out), gko, frame_origin(lam), frame_origin(phi), frame_origin(k), B,
                     X, XY, c, dt);
  int rc = check_hip(hipGetLastError(), kCtx);
  if (rc) return rc;
  return fill_ghosts_periodic(lam_out, stream);
}

}  // namespace
}  // namespace nebo

using namespace nebo;

extern "C" int nebo_diffusion_step(const nebo_field_t* out, const nebo_field_t* phi,
                                   const nebo_field_t* k, const nebo_axis_coefs_t* c, double dt,
                                   void* stream)
{
  clear_error();
  if (!out || !phi || !k || !c) return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  int rc;
  if ((rc = check_periodic_field(out, phi)) || (rc = check_periodic_field(phi, phi)) ||
      (rc = check_periodic_field(k, phi)))
    return rc;
  if (out->values == phi->values || out->values == k->values)
    return set_error(NEBO_ERR_INVALID, kCtx, "output must not alias an input");
  if ((rc = fill_ghosts_periodic(phi, stream))) return rc;
  return step_filled(out, phi, k, *c, dt, stream);
}

extern "C" int nebo_diffusion_adjoint_step(const nebo_field_t* lam_out, double* gk,
                                           const nebo_field_t* lam, const nebo_field_t* phi,
                                           const nebo_field_t* k, const nebo_axis_coefs_t* c,
                                           double dt, void* stream)
{
  clear_error();
  if (!lam_out || !gk || !lam || !phi || !k || !c) return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  int rc;
  if ((rc = check_periodic_field(lam_out, phi)) || (rc = check_periodic_field(lam, phi)) ||
      (rc = check_periodic_field(k, phi)) || (rc = check_periodic_field(phi, phi)))
    return rc;
  if (lam_out->values == lam->values || lam_out->values == phi->values || lam_out->values == k->values ||
      (void*)gk == lam_out->values)
    return set_error(NEBO_ERR_INVALID, kCtx, "outputs must not alias inputs");
  return adjoint_filled(lam_out, gk, lam, phi, k, *c, dt, stream);
}

extern "C" int nebo_diffusion_forward(const nebo_field_t* phiT, const nebo_field_t* phi0,
                                      const nebo_field_t* k, const nebo_axis_coefs_t* c, double dt,
                                      int32_t n_steps, int32_t every, double* ckpt, void* stream)
{
  clear_error();
  if (!phiT || !phi0 || !k || !c || !ckpt) return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  if (n_steps < 1 || every < 1 || n_steps % every)
    return set_error(NEBO_ERR_INVALID, kCtx, "n_steps must be a positive multiple of the checkpoint interval");
  int rc;
  if ((rc = check_periodic_field(phiT, phi0)) || (rc = check_periodic_field(phi0, phi0)) ||
      (rc = check_periodic_field(k, phi0)))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  const size_t np = field_npts(phi0), bytes = np * sizeof(double);
  double* tmp = nullptr;
  if ((rc = check_hip(hipMallocAsync((void**)&tmp, 2 * bytes, s), kCtx))) return rc;
  nebo_field_t ck0 = with_values(phi0, ckpt);
  rc = check_hip(hipMemcpyAsync(ckpt, phi0->values, bytes, hipMemcpyDeviceToDevice, s), kCtx);
  if (!rc) rc = fill_ghosts_periodic(&ck0, stream);
  double* cur = ckpt;
  for (int st = 0; st < n_steps && !rc; ++st) {
    double* dst;
    if ((st + 1) % every == 0) dst = ckpt + (size_t)((st + 1) / every) * np;
    else dst = (cur == tmp) ? tmp + np : tmp;
    nebo_field_t fin = with_values(phi0, cur), fout = with_values(phi0, dst);
    rc = step_filled(&fout, &fin, k, *c, dt, stream);
    cur = dst;
  }
  if (!rc && cur != phiT->values)
    rc = check_hip(hipMemcpyAsync(phiT->values, cur, bytes, hipMemcpyDeviceToDevice, s), kCtx);
  (void)hipFreeAsync(tmp, s);
  return rc;
}

extern "C" int nebo_diffusion_backward(const nebo_field_t* lam0, double* gk, const nebo_field_t* lamT,
                                       const nebo_field_t* k, const nebo_axis_coefs_t* c, double dt,
                                       int32_t n_steps, int32_t every, const double* ckpt,
                                       double* work, void* stream)
{
  clear_error();
  if (!lam0 || !gk || !lamT || !k || !c || !ckpt || !work)
    return set_error(NEBO_ERR_INVALID, kCtx, "null argument");
  if (n_steps < 1 || every < 1 || n_steps % every)
    return set_error(NEBO_ERR_INVALID, kCtx, "n_steps must be a positive multiple of the checkpoint interval");
  if (every > 1024) return set_error(NEBO_ERR_TOO_LARGE, kCtx, "checkpoint interval above 1024 steps");
  int rc;
  if ((rc = check_periodic_field(lam0, lamT)) || (rc = check_periodic_field(lamT, lamT)) ||
      (rc = check_periodic_field(k, lamT)))
    return rc;
  if (lam0->values == lamT->values) return set_error(NEBO_ERR_INVALID, kCtx, "lam0 must not alias lamT");
  hipStream_t s = (hipStream_t)stream;
  const size_t np = field_npts(lamT), bytes = np * sizeof(double);
  // lam ping-pongs between work slot `every` and lam0
  double* lbuf[2] = {work + (size_t)every * np, static_cast<double*>(lam0->values)};
  int cur = 0;
  nebo_field_t L0 = with_values(lamT, lbuf[0]);
  rc = check_hip(hipMemcpyAsync(lbuf[0], lamT->values, bytes, hipMemcpyDeviceToDevice, s), kCtx);
  if (!rc) rc = fill_ghosts_periodic(&L0, stream);
  if (!rc) rc = check_hip(hipMemsetAsync(gk, 0, field_npts(k) * sizeof(double), s), kCtx);
  const int nseg = n_steps / every;
  for (int seg = nseg - 1; seg >= 0 && !rc; --seg) {
    // recompute states every*seg ... every*seg + every - 1 (state 0 of the segment is the checkpoint)
    const double* states[1024];   // every <= 1024, checked on entry
    states[0] = ckpt + (size_t)seg * np;
    for (int t = 1; t < every && !rc; ++t) {
      nebo_field_t fin = with_values(lamT, const_cast<double*>(states[t - 1]));
      nebo_field_t fout = with_values(lamT, work + (size_t)(t - 1) * np);
      rc = step_filled(&fout, &fin, k, *c, dt, stream);
      states[t] = work + (size_t)(t - 1) * np;
    }
    for (int t = every - 1; t >= 0 && !rc; --t) {
      nebo_field_t lin = with_values(lamT, lbuf[cur]);
      nebo_field_t lout = with_values(lamT, lbuf[1 - cur]);
      nebo_field_t ph = with_values(lamT, const_cast<double*>(states[t]));
      rc = adjoint_filled(&lout, gk, &lin, &ph, k, *c, dt, stream);
      cur = 1 - cur;
    }
  }
  if (!rc && lbuf[cur] != lam0->values)
    rc = check_hip(hipMemcpyAsync(lam0->values, lbuf[cur], bytes, hipMemcpyDeviceToDevice, s), kCtx);
  return rc;
}
