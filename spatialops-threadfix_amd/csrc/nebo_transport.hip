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

struct Operands {
  const nebo_field_t* u;   // three face fields
  const nebo_field_t* k;
};

std::atomic<int> g_zchunk{0};   // planes per workgroup; 0: kZChunk

Box3 interior_box(const nebo_field_t* f)
{
  const int n[3] = {f->extent[0] - 2, f->extent[1] - 2, f->extent[2] - 2};
  return make_box(n, kTx, kTy, g_zchunk.load() > 0 ? g_zchunk.load() : kZChunk);
}

int launch_stage(const nebo_field_t* out, const nebo_field_t* pn, const nebo_field_t* s,
                 const Operands& o, const nebo_axis_coefs_t& c, double dt, int stage, void* stream)
{
  const Box3 B = interior_box(out);
  const int nb = B.nx * B.ny * B.nz;
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  auto fn = stage == 1 ? transport_stage_kernel<1> : stage == 2 ? transport_stage_kernel<2>
                                                                : transport_stage_kernel<3>;
  hipLaunchKernelGGL(fn, dim3(nb), dim3(kTx * kTy), 0, (hipStream_t)stream, frame_origin(out),
                     frame_origin(pn), frame_origin(stage == 1 ? pn : s), frame_origin(o.k),
                     frame_origin(&o.u[0]), frame_origin(&o.u[1]), frame_origin(&o.u[2]), B, X, XY,
                     coefs3(c), dt);
  int rc = check_hip(hipGetLastError(), kCtx);
  if (rc) return rc;
  return fill_ghosts_periodic(out, stream);
}

double stage_weight(int stage) { return stage == 3 ? 2.0 / 3.0 : stage == 2 ? 0.25 : 1.0; }

int launch_adjoint(const nebo_field_t* out, double* acc_out, double* gk, const nebo_field_t* a,
                   const double* acc_in, const nebo_field_t* s, const Operands& o,
                   const nebo_axis_coefs_t& c, double dt, int stage, void* stream)
{
  const Box3 B = interior_box(out);
  const int nb = B.nx * B.ny * B.nz;
  const int X = out->glob[0], XY = out->glob[0] * out->glob[1];
  // acc and gk are plain blocks with the field's shape: offset them like the field
  const std::ptrdiff_t off = frame_origin(out) - static_cast<double*>(out->values);
  const double cs = stage_weight(stage);
  const double cdt = cs * dt;
  auto fn = stage == 1 ? transport_adjoint_kernel<1> : stage == 2 ? transport_adjoint_kernel<2>
                                                                  : transport_adjoint_kernel<3>;
  hipLaunchKernelGGL(fn, dim3(nb), dim3(kTx * kTy), 0, (hipStream_t)stream, frame_origin(out),
                     acc_out ? acc_out + off : nullptr, gk + off, frame_origin(a),
                     acc_in ? acc_in + off : nullptr, frame_origin(s), frame_origin(o.k),
                     frame_origin(&o.u[0]), frame_origin(&o.u[1]), frame_origin(&o.u[2]), B, X, XY,
                     coefs3(c), dt, cs, cdt);
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

extern "C" int nebo_transport_set_zchunk(int32_t planes)
{
  clear_error();
  if (planes < 0 || planes > 4096)
    return set_error(NEBO_ERR_INVALID, kCtx, "z chunk: 0 (the default) or 1 to 4096 planes");
  g_zchunk.store(planes);
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
