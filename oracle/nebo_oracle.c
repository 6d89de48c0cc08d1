/*
 * nebo_oracle.c -- CPU restatement of the reference's Nebo evaluation semantics.
 *
 * TEST INFRASTRUCTURE ONLY (see nebo_oracle.h).  Built with -O2 -ffp-contract=off
 * so that every multiply and add rounds separately, as the reference's Release
 * build does without -march=native (CMakeLists.txt:18-22).
 *
 * What each part restates (reference paths relative to the repository root):
 *   ghost algebra ........ spatialops/structured/GhostData.h:54-257 (GHOST_MAX
 *                          "infinite" rule of operator+= at :176-187, min at :218-222,
 *                          point_to_ghost at :224-231)
 *   leaf/node ghosts ..... spatialops/NeboRhs.h:52-54,183-188; NeboStencils.h:311-317
 *                          with NeboStencilPointCollection::possible_ghosts :222-239;
 *                          NeboCond.h:86-91,284-289; NeboMask.h ghosts_with_bc;
 *                          NeboOperators.h (binary nodes take min of both operands)
 *   limits ............... spatialops/NeboBasic.h:61-100 and spatialops/NeboLhs.h:67-88
 *   serial walk .......... spatialops/NeboLhs.h:611-623 (z outer, x inner)
 *   leaf addressing ...... spatialops/NeboRhs.h:271-285 (base = values + window offset
 *                          + valid ghost minus); masks NeboMask.h:172-197 and
 *                          BitField.h:303-333 (window offset + total ghost minus)
 *   stencil fold ......... spatialops/NeboStencils.h:407-448 (left fold in point
 *                          definition order, value*coef); sum stencil :940-975;
 *                          edgeless stencil :526-760 (same fold, ghosts by
 *                          additive_point_to_ghost, GhostData.h:244-252)
 *   mask shift ........... spatialops/NeboStencils.h:1048-1213 (mask(x + P),
 *                          ghosts + additive_reductive_point_to_ghost(P))
 *   boundary conditions .. spatialops/NeboStencilBuilder.h:504-526 (mask shift
 *                          points) and :590-764 (cpu_apply: one pass over the
 *                          mask points, phi(g - shift) = (gamma - sum) / coef)
 *   particles ............ spatialops/particles/ParticleOperators.h:151-395
 *                          (serial particle loop, k/j/i cell loops, fractions)
 *   point fields ......... spatialops/pointfield/PointOperators.h:95-160
 *   pointwise ops ........ spatialops/NeboOperators.h (SeqWalk eval bodies, e.g.
 *                          :181, :923, :1665, :2407, :4303, :6668, :6875, :13237,
 *                          :13992, :14710)
 *   cond ................. spatialops/NeboCond.h:408-412 (first true clause wins)
 *   thread split ......... spatialops/NeboBasic.h:112-191, NeboLhs.h:335-393
 *   reductions ........... spatialops/NeboReductions.h:45-240
 *   masked assign ........ spatialops/NeboLhs.h:164-283
 *   bit field ............ spatialops/structured/BitField.h:291-362, SpatialMask.h:99-181
 *
 * Pinning (tests/test_oracle.py): every function above is checked against the
 * known answers of the reference's own tests (exact fused-vs-unfused chain
 * compare of test/StencilTest.cpp, cond/mask patterns of
 * spatialops/structured/test/testMask.cpp, the exact stencil-vs-hand-loop compare
 * of spatialops/structured/stencil/test/test_ref_stencil.cpp), against closed
 * forms (constant field => zero gradient, linear field => exact gradient and zero
 * Laplacian, periodic sum of a divergence == 0, second-order convergence of
 * div(k grad phi) on smooth fields), and against brute force on tiny grids.
 * The transcendental ops (sin ... erfc) use the host libm; erf/erfc of the
 * reference come from Boost.Math and are parity unpinned.
 */
#define _GNU_SOURCE
#include "nebo_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* ghost algebra                                                             */
/* ------------------------------------------------------------------------- */

static int g_is_inf(const oracle_ghost_t* g)
{
  for (int d = 0; d < 3; ++d)
    if (g->minus[d] < NEBO_GHOST_MAX || g->plus[d] < NEBO_GHOST_MAX) return 0;
  return 1;
}

static oracle_ghost_t g_uniform(int n)
{
  oracle_ghost_t g;
  for (int d = 0; d < 3; ++d) { g.minus[d] = n; g.plus[d] = n; }
  return g;
}

/* GhostData::operator+= : an infinite left side stays; an infinite right side wins. */
static oracle_ghost_t g_add(oracle_ghost_t a, oracle_ghost_t b)
{
  if (g_is_inf(&a)) return a;
  if (g_is_inf(&b)) return b;
  for (int d = 0; d < 3; ++d) { a.minus[d] += b.minus[d]; a.plus[d] += b.plus[d]; }
  return a;
}

static oracle_ghost_t g_sub(oracle_ghost_t a, oracle_ghost_t b)
{
  for (int d = 0; d < 3; ++d) { a.minus[d] -= b.minus[d]; a.plus[d] -= b.plus[d]; }
  return a;
}

static oracle_ghost_t g_min(oracle_ghost_t a, oracle_ghost_t b)
{
  for (int d = 0; d < 3; ++d) {
    if (b.minus[d] < a.minus[d]) a.minus[d] = b.minus[d];
    if (b.plus[d] < a.plus[d]) a.plus[d] = b.plus[d];
  }
  return a;
}

/* point_to_ghost(v): negative components go to minus, positive to plus. */
static oracle_ghost_t g_point(const int32_t v[3])
{
  oracle_ghost_t g;
  for (int d = 0; d < 3; ++d) {
    g.minus[d] = v[d] < 0 ? -v[d] : 0;
    g.plus[d] = v[d] > 0 ? v[d] : 0;
  }
  return g;
}

/* possible_ghosts of a point list: additive_reductive_point_to_ghost(p) is
 * (minus = p, plus = -p) per axis, folded with min over the points. */
static oracle_ghost_t g_stencil_shift(const nebo_stencil_t* s)
{
  oracle_ghost_t g = g_uniform(0);
  for (int i = 0; i < s->npts; ++i)
    for (int d = 0; d < 3; ++d) {
      int mn = s->pts[i][d], pl = -s->pts[i][d];
      if (i == 0 || mn < g.minus[d]) g.minus[d] = mn;
      if (i == 0 || pl < g.plus[d]) g.plus[d] = pl;
    }
  return g;
}

/* Edgeless stencils: additive_point_to_ghost(p) = (minus max(p,0), plus max(-p,0)),
 * folded with min over the points (the reference compiles the one-point case). */
static oracle_ghost_t g_edgeless_shift(const nebo_stencil_t* s)
{
  oracle_ghost_t g = g_uniform(0);
  for (int i = 0; i < s->npts; ++i)
    for (int d = 0; d < 3; ++d) {
      int v = s->pts[i][d];
      int mn = v > 0 ? v : 0, pl = v < 0 ? -v : 0;
      if (i == 0 || mn < g.minus[d]) g.minus[d] = mn;
      if (i == 0 || pl < g.plus[d]) g.plus[d] = pl;
    }
  return g;
}

int oracle_node_ghosts(const nebo_expr_t* e, int node, int with_bc, oracle_ghost_t* out)
{
  if (node < 0 || node >= e->n_nodes) return NEBO_ERR_INVALID;
  const nebo_node_t* n = &e->nodes[node];
  oracle_ghost_t ga, gb, gc;
  int rc;
  switch (n->op) {
  case NEBO_SCALAR:
    *out = g_uniform(NEBO_GHOST_MAX);
    return NEBO_OK;
  case NEBO_FIELD: {
    const nebo_field_t* f = &e->fields[n->idx];
    oracle_ghost_t g;
    for (int d = 0; d < 3; ++d) { g.minus[d] = f->valid_minus[d]; g.plus[d] = f->valid_plus[d]; }
    if (with_bc) g = g_add(g, g_point(f->bc_extra));
    *out = g;
    return NEBO_OK;
  }
  case NEBO_MASK: {
    const nebo_mask_t* m = &e->masks[n->idx];
    oracle_ghost_t g;
    for (int d = 0; d < 3; ++d) { g.minus[d] = m->valid_minus[d]; g.plus[d] = m->valid_plus[d]; }
    if (with_bc) g = g_add(g, g_point(m->bc_extra));
    *out = g;
    return NEBO_OK;
  }
  case NEBO_STENCIL: {
    const nebo_stencil_t* s = &e->stencils[n->idx];
    if ((rc = oracle_node_ghosts(e, n->a, with_bc, &ga))) return rc;
    *out = g_add(ga, s->kind == NEBO_STENCIL_EDGELESS ? g_edgeless_shift(s) : g_stencil_shift(s));
    return NEBO_OK;
  }
  case NEBO_MASK_SHIFT:
    if ((rc = oracle_node_ghosts(e, n->a, with_bc, &ga))) return rc;
    *out = g_add(ga, g_stencil_shift(&e->stencils[n->idx]));
    return NEBO_OK;
  case NEBO_COND:
    if ((rc = oracle_node_ghosts(e, n->a, with_bc, &ga))) return rc;
    if ((rc = oracle_node_ghosts(e, n->b, with_bc, &gb))) return rc;
    if ((rc = oracle_node_ghosts(e, n->c, with_bc, &gc))) return rc;
    *out = g_min(g_min(ga, gb), gc);
    return NEBO_OK;
  case NEBO_NEG: case NEBO_ABS: case NEBO_SQRT: case NEBO_SQUARE: case NEBO_CUBE:
  case NEBO_POWINT: case NEBO_EXP: case NEBO_LOG: case NEBO_LOG10: case NEBO_SIN:
  case NEBO_COS: case NEBO_TAN: case NEBO_TANH: case NEBO_ERF: case NEBO_ERFC:
  case NEBO_INV_ERF: case NEBO_INV_ERFC: case NEBO_NOT:
    return oracle_node_ghosts(e, n->a, with_bc, out);
  default:
    if (n->op < 0 || n->op >= NEBO_NUM_OPCODES) return NEBO_ERR_INVALID;
    if ((rc = oracle_node_ghosts(e, n->a, with_bc, &ga))) return rc;
    if ((rc = oracle_node_ghosts(e, n->b, with_bc, &gb))) return rc;
    *out = g_min(ga, gb);
    return NEBO_OK;
  }
}

/* has_extents()/extents() of the tree (first operand with extents wins). */
static int node_extents(const nebo_expr_t* e, int node, int32_t ext[3])
{
  const nebo_node_t* n = &e->nodes[node];
  switch (n->op) {
  case NEBO_SCALAR: return 0;
  case NEBO_FIELD: {
    const nebo_field_t* f = &e->fields[n->idx];
    for (int d = 0; d < 3; ++d) ext[d] = f->extent[d] - f->valid_minus[d] - f->valid_plus[d];
    return 1;
  }
  case NEBO_MASK: {
    const nebo_mask_t* m = &e->masks[n->idx];
    for (int d = 0; d < 3; ++d) ext[d] = m->extent[d] - m->ghost_minus[d] - m->ghost_plus[d];
    return 1;
  }
  case NEBO_STENCIL: case NEBO_MASK_SHIFT: return node_extents(e, n->a, ext);
  case NEBO_COND:
    if (node_extents(e, n->a, ext)) return 1;
    if (node_extents(e, n->b, ext)) return 1;
    return node_extents(e, n->c, ext);
  case NEBO_NEG: case NEBO_ABS: case NEBO_SQRT: case NEBO_SQUARE: case NEBO_CUBE:
  case NEBO_POWINT: case NEBO_EXP: case NEBO_LOG: case NEBO_LOG10: case NEBO_SIN:
  case NEBO_COS: case NEBO_TAN: case NEBO_TANH: case NEBO_ERF: case NEBO_ERFC:
  case NEBO_INV_ERF: case NEBO_INV_ERFC: case NEBO_NOT:
    return node_extents(e, n->a, ext);
  default:
    if (node_extents(e, n->a, ext)) return 1;
    return node_extents(e, n->b, ext);
  }
}

int oracle_assign_limits(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                         nebo_limits_t* lim, int* empty)
{
  oracle_ghost_t r;
  int rc = oracle_node_ghosts(rhs, rhs->n_nodes - 1, 1, &r);
  if (rc) return rc;
  /* calculate_actual_ghost */
  for (int d = 0; d < 3; ++d)
    if (lhs->has_bc[d] && r.plus[d] < lhs->bc_extra[d]) return NEBO_ERR_GHOST;
  oracle_ghost_t extra = g_point(lhs->bc_extra);
  oracle_ghost_t g;
  if (use_ghost) {
    oracle_ghost_t l;
    for (int d = 0; d < 3; ++d) { l.minus[d] = lhs->ghost_minus[d]; l.plus[d] = lhs->ghost_plus[d]; }
    g = g_sub(g_min(g_add(l, extra), r), extra);
  } else {
    oracle_ghost_t l = g_uniform(0);
    for (int d = 0; d < 3; ++d) l.plus[d] = lhs->bc_extra[d];
    g = g_sub(l, extra);
  }
  int vol = 1;
  for (int d = 0; d < 3; ++d) {
    int ext = lhs->extent[d] - lhs->valid_minus[d] - lhs->valid_plus[d];
    lim->lo[d] = -g.minus[d];
    lim->hi[d] = ext + g.plus[d];
    if (lim->hi[d] - lim->lo[d] <= 0) vol = 0;
  }
  *empty = !vol;
  return NEBO_OK;
}

/* ------------------------------------------------------------------------- */
/* tree walk                                                                 */
/* ------------------------------------------------------------------------- */

static inline size_t field_index(const nebo_field_t* f, int x, int y, int z)
{
  long gx = f->offset[0] + f->valid_minus[0] + x;
  long gy = f->offset[1] + f->valid_minus[1] + y;
  long gz = f->offset[2] + f->valid_minus[2] + z;
  return (size_t)(gx + (long)f->glob[0] * (gy + (long)f->glob[1] * gz));
}

static inline int mask_bit(const nebo_mask_t* m, int x, int y, int z)
{
  long gx = m->offset[0] + m->ghost_minus[0] + x;
  long gy = m->offset[1] + m->ghost_minus[1] + y;
  long gz = m->offset[2] + m->ghost_minus[2] + z;
  long pos = gx + (long)m->glob[0] * (gy + (long)m->glob[1] * gz);
  return (m->bits[pos / 32] >> (pos % 32)) & 1u;
}

/* erf^-1 and erfc^-1: boost::math::erf_inv / erfc_inv in the reference's CPU
 * walk (NeboOperators.h:6252-6262 and the InvErfcFcn that follows; tested at 5
 * and 6 ulp in test/NeboTest.cpp:521-523).  Boost.Math is not available here, so
 * they are restated from libm's erf and erfc: a start from Winitzki's closed-form
 * approximation (S. Winitzki, "A handy approximation for the error function and
 * its inverse", 2008; relative error below 2e-3) refined by Newton steps on the
 * equation that keeps the residual accurate: erf(x) = y where x is small
 * (|y| <= 0.5), ln(erfc(x)/q) = 0 in the tails.  End points give +-inf and
 * arguments outside the domain NaN, as the GPU branch's erfinv/erfcinv do
 * (Boost's default policy throws there instead). */
static double winitzki_start(double L)   /* L = ln(1 - y^2); returns |x| */
{
  const double a = 0.147;
  const double t = 2.0 / (M_PI * a) + 0.5 * L;
  const double r = sqrt(t * t - L / a) - t;
  return r > 0.0 ? sqrt(r) : 0.0;
}

/* e^{-x^2} / erfc(x) for x >= 0, used only as the Newton slope (its accuracy sets
 * the convergence rate, not the root): exact form while e^{-x^2} is a normal
 * number, the asymptotic series of erfc beyond. */
static double erfc_slope_ratio(double x)
{
  if (x < 26.0) return exp(-x * x) / erfc(x);
  const double u = 1.0 / (2.0 * x * x);
  return sqrt(M_PI) * x / (1.0 - u + 3.0 * u * u - 15.0 * u * u * u);
}

/* ln(erfc(x)/q) for x >= 0: through erfc while erfc(x) is a normal number,
 * else ln erfc(x) from its asymptotic series
 *   erfc(x) = e^{-x^2} / (x sqrt(pi)) * sum_k (-1)^k (2k-1)!! / (2x^2)^k
 * (10 terms; at x >= 26 the last one is below 1e-30), so a denormal q still has
 * an accurate residual. */
static double tail_residual(double x, double q)
{
  if (x < 26.0) return log1p((erfc(x) - q) / q);
  const double u = 1.0 / (2.0 * x * x);
  double term = 1.0, sum = 1.0;
  for (int k = 1; k <= 10; ++k) {
    term *= -(2.0 * k - 1.0) * u;
    sum += term;
  }
  return (-x * x - log(x) - 0.5 * log(M_PI) + log(sum)) - log(q);
}

/* x >= 0 with erfc(x) = q, 0 < q <= 0.5 */
static double inv_erfc_tail(double q)
{
  double x = winitzki_start(log(q * (2.0 - q)));
  for (int it = 0; it < 8; ++it) {
    const double h = tail_residual(x, q);
    const double hp = -(2.0 / sqrt(M_PI)) * erfc_slope_ratio(x);
    double xn = x - h / hp;
    if (xn > 28.0) xn = 28.0;   /* erfc_inv of the least denormal is 27.21 */
    if (xn < 0.0) xn = 0.0;
    x = xn;
  }
  return x;
}

/* x with erf(x) = y, |y| <= 0.5 */
static double inv_erf_centre(double y)
{
  double x = copysign(winitzki_start(log1p(-y * y)), y);
  for (int it = 0; it < 6; ++it) x = x - (erf(x) - y) / ((2.0 / sqrt(M_PI)) * exp(-x * x));
  return x;
}

double oracle_inv_erf(double y)
{
  if (isnan(y) || fabs(y) > 1.0) return NAN;
  if (y == 0.0) return y;
  if (fabs(y) == 1.0) return copysign(INFINITY, y);
  if (fabs(y) <= 0.5) return inv_erf_centre(y);
  return copysign(inv_erfc_tail(1.0 - fabs(y)), y);   /* 1 - |y| is exact here */
}

double oracle_inv_erfc(double q)
{
  if (isnan(q) || q < 0.0 || q > 2.0) return NAN;
  if (q == 0.0) return INFINITY;
  if (q == 2.0) return -INFINITY;
  if (q > 1.0) return -oracle_inv_erfc(2.0 - q);   /* exact: 1 < q < 2 */
  if (q >= 0.5) return q == 1.0 ? 0.0 : inv_erf_centre(1.0 - q);   /* erf(x) = 1 - q, exact */
  return inv_erfc_tail(q);
}

double oracle_eval(const nebo_expr_t* e, int node, int x, int y, int z)
{
  const nebo_node_t* n = &e->nodes[node];
  double a, b;
  switch (n->op) {
  case NEBO_SCALAR: return n->value;
  case NEBO_FIELD: {
    /* reads outside the allocation (only an edgeless stencil can make them, at
       points a cond does not select) give NaN instead of touching memory */
    const nebo_field_t* f = &e->fields[n->idx];
    long i = (long)f->offset[0] + f->valid_minus[0] + x +
             (long)f->glob[0] * ((long)f->offset[1] + f->valid_minus[1] + y +
                                 (long)f->glob[1] * ((long)f->offset[2] + f->valid_minus[2] + z));
    if (i < 0 || i >= (long)f->glob[0] * f->glob[1] * f->glob[2]) return NAN;
    return ((const double*)f->values)[i];
  }
  case NEBO_MASK: return mask_bit(&e->masks[n->idx], x, y, z) ? 1.0 : 0.0;
  case NEBO_MASK_SHIFT: {
    const int32_t* P = e->stencils[n->idx].pts[0];
    return mask_bit(&e->masks[e->nodes[n->a].idx], x + P[0], y + P[1], z + P[2]) ? 1.0 : 0.0;
  }
  case NEBO_ADD: return oracle_eval(e, n->a, x, y, z) + oracle_eval(e, n->b, x, y, z);
  case NEBO_SUB: return oracle_eval(e, n->a, x, y, z) - oracle_eval(e, n->b, x, y, z);
  case NEBO_MUL: return oracle_eval(e, n->a, x, y, z) * oracle_eval(e, n->b, x, y, z);
  case NEBO_DIV: return oracle_eval(e, n->a, x, y, z) / oracle_eval(e, n->b, x, y, z);
  case NEBO_NEG: return -(oracle_eval(e, n->a, x, y, z));
  case NEBO_MAX:
    a = oracle_eval(e, n->a, x, y, z); b = oracle_eval(e, n->b, x, y, z);
    return (a > b) ? a : b;
  case NEBO_MIN:
    a = oracle_eval(e, n->a, x, y, z); b = oracle_eval(e, n->b, x, y, z);
    return (a < b) ? a : b;
  case NEBO_ABS: return fabs(oracle_eval(e, n->a, x, y, z));
  case NEBO_SQRT: return sqrt(oracle_eval(e, n->a, x, y, z));
  case NEBO_SQUARE: a = oracle_eval(e, n->a, x, y, z); return a * a;
  case NEBO_CUBE: a = oracle_eval(e, n->a, x, y, z); return a * a * a;
  case NEBO_POW: return pow(oracle_eval(e, n->a, x, y, z), oracle_eval(e, n->b, x, y, z));
  case NEBO_POWINT: return pow(oracle_eval(e, n->a, x, y, z), (double)n->idx);
  case NEBO_EXP: return exp(oracle_eval(e, n->a, x, y, z));
  case NEBO_LOG: return log(oracle_eval(e, n->a, x, y, z));
  case NEBO_LOG10: return log10(oracle_eval(e, n->a, x, y, z));
  case NEBO_SIN: return sin(oracle_eval(e, n->a, x, y, z));
  case NEBO_COS: return cos(oracle_eval(e, n->a, x, y, z));
  case NEBO_TAN: return tan(oracle_eval(e, n->a, x, y, z));
  case NEBO_TANH: return tanh(oracle_eval(e, n->a, x, y, z));
  case NEBO_ERF: return erf(oracle_eval(e, n->a, x, y, z));
  case NEBO_ERFC: return erfc(oracle_eval(e, n->a, x, y, z));
  case NEBO_INV_ERF: return oracle_inv_erf(oracle_eval(e, n->a, x, y, z));
  case NEBO_INV_ERFC: return oracle_inv_erfc(oracle_eval(e, n->a, x, y, z));
  case NEBO_EQ: return oracle_eval(e, n->a, x, y, z) == oracle_eval(e, n->b, x, y, z);
  case NEBO_NE: return oracle_eval(e, n->a, x, y, z) != oracle_eval(e, n->b, x, y, z);
  case NEBO_LT: return oracle_eval(e, n->a, x, y, z) < oracle_eval(e, n->b, x, y, z);
  case NEBO_LE: return oracle_eval(e, n->a, x, y, z) <= oracle_eval(e, n->b, x, y, z);
  case NEBO_GT: return oracle_eval(e, n->a, x, y, z) > oracle_eval(e, n->b, x, y, z);
  case NEBO_GE: return oracle_eval(e, n->a, x, y, z) >= oracle_eval(e, n->b, x, y, z);
  case NEBO_AND:
    return (oracle_eval(e, n->a, x, y, z) != 0.0) && (oracle_eval(e, n->b, x, y, z) != 0.0);
  case NEBO_OR:
    return (oracle_eval(e, n->a, x, y, z) != 0.0) || (oracle_eval(e, n->b, x, y, z) != 0.0);
  case NEBO_NOT: return !(oracle_eval(e, n->a, x, y, z) != 0.0);
  case NEBO_STENCIL: {
    const nebo_stencil_t* s = &e->stencils[n->idx];
    double acc = 0.0;
    for (int i = 0; i < s->npts; ++i) {
      double v = oracle_eval(e, n->a, x + s->pts[i][0], y + s->pts[i][1], z + s->pts[i][2]);
      if (s->kind != NEBO_STENCIL_SUM) v = v * s->coefs[i];
      acc = (i == 0) ? v : acc + v;
    }
    return acc;
  }
  case NEBO_COND:
    return (oracle_eval(e, n->a, x, y, z) != 0.0) ? oracle_eval(e, n->b, x, y, z)
                                                  : oracle_eval(e, n->c, x, y, z);
  default: return NAN;
  }
}

static void walk(const nebo_field_t* lhs, const nebo_expr_t* rhs, const nebo_limits_t* L)
{
  double* out = (double*)lhs->values;
  int root = rhs->n_nodes - 1;
  for (int z = L->lo[2]; z < L->hi[2]; ++z)
    for (int y = L->lo[1]; y < L->hi[1]; ++y)
      for (int x = L->lo[0]; x < L->hi[0]; ++x)
        out[field_index(lhs, x, y, z)] = oracle_eval(rhs, root, x, y, z);
}

int oracle_assign(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost)
{
  nebo_limits_t L;
  int empty;
  int rc = oracle_assign_limits(lhs, rhs, use_ghost, &L, &empty);
  if (rc) return rc;
  if (!empty) walk(lhs, rhs, &L);
  return NEBO_OK;
}

/* ------------------------------------------------------------------------- */
/* thread-parallel split                                                     */
/* ------------------------------------------------------------------------- */

/* nebo_find_partition: split only one axis, z if possible, else y, else x. */
static void find_partition(const int32_t ext[3], int nthreads, int32_t split[3])
{
  split[0] = split[1] = split[2] = 1;
  if (nthreads <= ext[2]) split[2] = nthreads;
  else if (nthreads <= ext[1]) split[1] = nthreads;
  else if (nthreads <= ext[0]) split[0] = nthreads;
}

/* nebo_set_up_extents: balanced pieces, the first (full % split) get one more. */
static void set_up_extents(const int32_t cur[3], const int32_t split[3],
                           const nebo_limits_t* L, nebo_limits_t* out)
{
  for (int d = 0; d < 3; ++d) {
    int full = L->hi[d] - L->lo[d];
    int std = full / split[d];
    int nextra = full % split[d];
    int past = cur[d] < nextra ? cur[d] : nextra;
    int mine = cur[d] < nextra ? 1 : 0;
    int lo = std * cur[d] + past;
    out->lo[d] = lo + L->lo[d];
    out->hi[d] = lo + std + mine + L->lo[d];
  }
}

/* nebo_next_partition: z fastest, then y, then x. */
static void next_partition(int32_t cur[3], const int32_t split[3])
{
  if (cur[2] < split[2] - 1) cur[2]++;
  else if (cur[1] < split[1] - 1) { cur[1]++; cur[2] = 0; }
  else { cur[0]++; cur[1] = 0; cur[2] = 0; }
}

typedef struct task_arg {
  const nebo_field_t* lhs;
  const nebo_expr_t* rhs;
  nebo_limits_t lim;
  /* closed-form variant */
  const nebo_field_t *k, *phi;
  const nebo_axis_coefs_t* c;
  int form;
} task_arg_t;

static void* task_walk(void* p)
{
  task_arg_t* t = (task_arg_t*)p;
  walk(t->lhs, t->rhs, &t->lim);
  return NULL;
}

static void run_tasks(task_arg_t* tasks, int ntasks, void* (*fn)(void*))
{
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)ntasks);
  /* task 0 runs on the caller, like Threading::scheduleTasks (Threading.h:25-37) */
  for (int i = 1; i < ntasks; ++i) pthread_create(&th[i], NULL, fn, &tasks[i]);
  fn(&tasks[0]);
  for (int i = 1; i < ntasks; ++i) pthread_join(th[i], NULL);
  free(th);
}

static int make_tasks(const nebo_limits_t* L, int nthreads, task_arg_t** out)
{
  int32_t ext[3] = {L->hi[0] - L->lo[0], L->hi[1] - L->lo[1], L->hi[2] - L->lo[2]};
  int32_t split[3];
  find_partition(ext, nthreads, split);
  int n = split[0] * split[1] * split[2];
  task_arg_t* t = (task_arg_t*)calloc((size_t)n, sizeof(task_arg_t));
  int32_t cur[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i) {
    set_up_extents(cur, split, L, &t[i].lim);
    next_partition(cur, split);
  }
  *out = t;
  return n;
}

int oracle_assign_threads(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                          int nthreads)
{
  nebo_limits_t L;
  int empty;
  int rc = oracle_assign_limits(lhs, rhs, use_ghost, &L, &empty);
  if (rc) return rc;
  if (empty) return NEBO_OK;
  task_arg_t* t;
  int n = make_tasks(&L, nthreads, &t);
  for (int i = 0; i < n; ++i) { t[i].lhs = lhs; t[i].rhs = rhs; }
  run_tasks(t, n, task_walk);
  free(t);
  return NEBO_OK;
}

int oracle_masked_assign(const nebo_field_t* lhs, const int32_t* pts, int n,
                         const nebo_expr_t* rhs, int passes)
{
  double* out = (double*)lhs->values;
  int root = rhs->n_nodes - 1;
  for (int p = 0; p < passes; ++p)
    for (int i = 0; i < n; ++i) {
      int x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
      out[field_index(lhs, x, y, z)] = oracle_eval(rhs, root, x, y, z);
    }
  return NEBO_OK;
}

/* Boundary-condition inversion on the CPU (NeboBoundaryConditionBuilder,
 * NeboStencilBuilder.h:590-764, shift points :504-526).  The operator's points
 * are in definition order: the low point is the first defined (the reference's
 * PointCollection::Last, coefficient coefs().last()), the high point the last
 * (PointCollection::First, coefs().coef()).  Minus side: shift = (gamma offset <
 * phi offset) per component, or e_axis for a same-location (FD) operator; plus
 * side: shift = -(gamma offset > phi offset), or -e_axis.  For each mask point g
 * (gamma frame) in list order, x = g - shift and
 *   phi(x) = (gamma(x - p_side) * 1.0 - fold_{i != side} phi(x + p_i - p_side) * c_i) / c_side
 * with the fold left to right in definition order, written at once (cpu_apply
 * runs a single pass). */
int oracle_bc_apply(const nebo_field_t* phi, const nebo_expr_t* gamma,
                    const nebo_stencil_t* op, const int32_t phi_off[3],
                    const int32_t gamma_off[3], int fd_axis, const int32_t* pts, int n,
                    int minus)
{
  if (op->npts < 2) return NEBO_ERR_INVALID;
  int side = minus ? 0 : op->npts - 1;
  int32_t shift[3];
  for (int d = 0; d < 3; ++d) {
    if (fd_axis >= 0)
      shift[d] = (d == fd_axis) ? (minus ? 1 : -1) : 0;
    else
      shift[d] = minus ? (gamma_off[d] < phi_off[d]) : -(gamma_off[d] > phi_off[d]);
  }
  const int32_t* ps = op->pts[side];
  double* out = (double*)phi->values;
  int root = gamma->n_nodes - 1;
  for (int i = 0; i < n; ++i) {
    int x = pts[3 * i] - shift[0], y = pts[3 * i + 1] - shift[1], z = pts[3 * i + 2] - shift[2];
    double g = oracle_eval(gamma, root, x - ps[0], y - ps[1], z - ps[2]) * 1.0;
    double acc = 0.0;
    int first = 1;
    for (int k = 0; k < op->npts; ++k) {
      if (k == side) continue;
      int qx = x + op->pts[k][0] - ps[0], qy = y + op->pts[k][1] - ps[1],
          qz = z + op->pts[k][2] - ps[2];
      double v = out[field_index(phi, qx, qy, qz)] * op->coefs[k];
      acc = first ? v : acc + v;
      first = 0;
    }
    out[field_index(phi, x, y, z)] = (g - acc) / op->coefs[side];
  }
  return NEBO_OK;
}

/* ParticleToCell / CellToParticle, restated from ParticleOperators.h:151-395.
 * The reference converts the low cell index to size_t: exact (truncation) for
 * t > -1; for t <= -1 the conversion is undefined and wraps in practice, which
 * leaves the particle's loops empty.  The restatement takes that practical
 * result: such a particle touches no cell. */
typedef struct {
  double h[3], face[3];
  int has[3];
  long lo[3], hi[3];
  double plo[3], phi[3], pvol;
  int skip;
} pbox_t;

static void particle_setup(pbox_t* b, const nebo_field_t* f, const nebo_particle_geom_t* g,
                           const double* const c[3], const double* psize, long p)
{
  const double d[3] = {g->dx, g->dy, g->dz}, o[3] = {g->xlo, g->ylo, g->zlo};
  const double size = psize[p], rp = size * 0.5;
  b->skip = 0;
  for (int a = 0; a < 3; ++a) {
    /* constructor: dx_ = dx<=0 ? 1 : dx, xlo_ = dx<=0 ? 0 : xlo (:151-166) */
    double h = d[a] <= 0 ? 1.0 : d[a], lo = d[a] <= 0 ? 0.0 : o[a];
    b->h[a] = h;
    /* xloface = xlo_ - dx_*ngm[0] - dx_/2 (:190-192) */
    b->face[a] = lo - h * f->ghost_minus[a] - h / 2;
    b->has[a] = c[a] != NULL;
    b->plo[a] = b->has[a] ? c[a][p] - rp : 0;
    b->phi[a] = b->has[a] ? b->plo[a] + size : 0;
    double t = (b->plo[a] - b->face[a]) / h;
    if (!(t > -1.0)) { b->skip = 1; t = 0; }
    b->lo[a] = (long)t;
    if (b->has[a]) {
      double u = (b->phi[a] - b->face[a]) / h;
      long ui = (long)u + 1;   /* int(u) + 1, then std::min with the window extent */
      b->hi[a] = ui < f->extent[a] ? ui : f->extent[a];
    } else {
      b->hi[a] = b->lo[a] + 1;
      if (b->hi[a] > f->extent[a]) b->hi[a] = f->extent[a];
    }
  }
  b->pvol = (b->has[0] ? 2 * rp : 1) * (b->has[1] ? 2 * rp : 1) * (b->has[2] ? 2 * rp : 1);
}

/* dest(i,j,k) of a window index (MemoryWindow::flat_index, MemoryWindow.h:121-131) */
static size_t window_index(const nebo_field_t* f, long i, long j, long k)
{
  long gi = f->glob[0] > 1 ? i + f->offset[0] : 0;
  long gj = f->glob[1] > 1 ? j + f->offset[1] : 0;
  long gk = f->glob[2] > 1 ? k + f->offset[2] : 0;
  return (size_t)(gi + (long)f->glob[0] * (gj + (long)f->glob[1] * gk));
}

static double pmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static double pmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* Sum (C2P, dest_cell == NULL) or scatter (P2C) one particle's fractions. */
static double particle_pass(const pbox_t* b, const nebo_field_t* f, double* cells, double value)
{
  double acc = 0.0;
  if (b->skip) return 0.0;
  for (long k = b->lo[2]; k < b->hi[2]; ++k) {
    double zcm = b->face[2] + k * b->h[2], zcp = zcm + b->h[2];
    double zcont = b->has[2] ? pmin(zcp, b->phi[2]) - pmax(zcm, b->plo[2]) : 1;
    for (long j = b->lo[1]; j < b->hi[1]; ++j) {
      double ycm = b->face[1] + j * b->h[1], ycp = ycm + b->h[1];
      double ycont = b->has[1] ? pmin(ycp, b->phi[1]) - pmax(ycm, b->plo[1]) : 1;
      for (long i = b->lo[0]; i < b->hi[0]; ++i) {
        double xcm = b->face[0] + i * b->h[0], xcp = xcm + b->h[0];
        double xcont = b->has[0] ? pmin(xcp, b->phi[0]) - pmax(xcm, b->plo[0]) : 1;
        double contribution = xcont * ycont * zcont / b->pvol;
        size_t at = window_index(f, i, j, k);
        if (cells == NULL)
          acc += ((const double*)f->values)[at] * contribution;
        else
          cells[at] += value * contribution;
      }
    }
  }
  return acc;
}

int oracle_particle_to_cell(const nebo_field_t* dest, const double* src, const double* px,
                            const double* py, const double* pz, const double* psize, int64_t np,
                            const nebo_particle_geom_t* geom)
{
  /* dest <<= 0.0 */
  nebo_node_t zero;
  memset(&zero, 0, sizeof zero);
  zero.op = NEBO_SCALAR;
  nebo_expr_t e = {1, &zero, 0, NULL, 0, NULL, 0, NULL};
  int rc = oracle_assign(dest, &e, 1);
  if (rc) return rc;
  const double* c[3] = {px, py, pz};
  for (long p = 0; p < np; ++p) {
    pbox_t b;
    particle_setup(&b, dest, geom, c, psize, p);
    particle_pass(&b, dest, (double*)dest->values, src[p]);
  }
  return NEBO_OK;
}

int oracle_cell_to_particle(double* dest, const nebo_field_t* src, const double* px,
                            const double* py, const double* pz, const double* psize, int64_t np,
                            const nebo_particle_geom_t* geom)
{
  const double* c[3] = {px, py, pz};
  for (long p = 0; p < np; ++p) {
    pbox_t b;
    particle_setup(&b, src, geom, c, psize, p);
    dest[p] = particle_pass(&b, src, NULL, 0.0);
  }
  return NEBO_OK;
}

/* PointToField / FieldToPoint (PointOperators.h:95-160): an in-order loop. */
void oracle_point_to_field(double* field_values, const double* src, const int64_t* idx, int64_t n)
{
  for (int64_t i = 0; i < n; ++i) field_values[idx[i]] = src[i];
}

void oracle_field_to_point(double* dest, const double* field_values, const int64_t* idx, int64_t n)
{
  for (int64_t i = 0; i < n; ++i) dest[i] = field_values[idx[i]];
}

/* ------------------------------------------------------------------------- */
/* reductions                                                                */
/* ------------------------------------------------------------------------- */

static int fold_limits(const nebo_expr_t* e, int use_ghost, nebo_limits_t* L)
{
  int root = e->n_nodes - 1;
  oracle_ghost_t g = g_uniform(0);
  if (use_ghost) {
    int rc = oracle_node_ghosts(e, root, 0, &g);
    if (rc) return rc;
    /* no field and no mask: "infinite" ghosts; the reference's loop would walk
     * [-GHOST_MAX, GHOST_MAX)^3 (5.8e12 cells).  The engine refuses such a
     * reduction, and so does the checker rather than run for hours. */
    for (int d = 0; d < 3; ++d)
      if (g.minus[d] >= NEBO_GHOST_MAX || g.plus[d] >= NEBO_GHOST_MAX) return NEBO_ERR_INVALID;
  }
  int32_t ext[3] = {0, 0, 0};
  node_extents(e, root, ext);
  for (int d = 0; d < 3; ++d) { L->lo[d] = -g.minus[d]; L->hi[d] = ext[d] + g.plus[d]; }
  return NEBO_OK;
}

int oracle_reduce(const nebo_expr_t* e, int op, int use_ghost, double* out)
{
  nebo_limits_t L;
  int rc = fold_limits(e, use_ghost, &L);
  if (rc) return rc;
  int root = e->n_nodes - 1;
  double r = op == NEBO_RED_MAX ? -INFINITY : op == NEBO_RED_MIN ? INFINITY : 0.0;
  for (int z = L.lo[2]; z < L.hi[2]; ++z)
    for (int y = L.lo[1]; y < L.hi[1]; ++y)
      for (int x = L.lo[0]; x < L.hi[0]; ++x) {
        double v = oracle_eval(e, root, x, y, z);
        switch (op) {
        case NEBO_RED_MAX: r = (r < v) ? v : r; break;
        case NEBO_RED_MIN: r = (r > v) ? v : r; break;
        case NEBO_RED_SUM: r = r + v; break;
        case NEBO_RED_NORM: r = r + pow(v, 2.0); break;
        default: return NEBO_ERR_INVALID;
        }
      }
  *out = op == NEBO_RED_NORM ? sqrt(r) : r;
  return NEBO_OK;
}

/* ------------------------------------------------------------------------- */
/* ghosts and masks                                                          */
/* ------------------------------------------------------------------------- */

void oracle_fill_ghosts_periodic(const nebo_field_t* f)
{
  double* v = (double*)f->values;
  const long X = f->glob[0], Y = f->glob[1];
  for (int d = 0; d < 3; ++d) {
    int gm = f->ghost_minus[d], gp = f->ghost_plus[d];
    int n = f->extent[d] - gm - gp;
    if (n <= 0 || (gm == 0 && gp == 0)) continue;
    int o = f->offset[d];
    /* the two other axes run over their whole window (ghosts included) */
    int a0 = (d + 1) % 3, a1 = (d + 2) % 3;
    for (int j = 0; j < f->extent[a1]; ++j)
      for (int i = 0; i < f->extent[a0]; ++i) {
        int idx[3];
        idx[a0] = f->offset[a0] + i;
        idx[a1] = f->offset[a1] + j;
        for (int g = 0; g < gm; ++g) { /* minus ghost g+1 layers out <- interior from the top */
          int dst = o + gm - 1 - g, src = o + gm + n - 1 - g;
          idx[d] = dst; long di = idx[0] + X * (idx[1] + Y * (long)idx[2]);
          idx[d] = src; long si = idx[0] + X * (idx[1] + Y * (long)idx[2]);
          v[di] = v[si];
        }
        for (int g = 0; g < gp; ++g) {
          int dst = o + gm + n + g, src = o + gm + g;
          idx[d] = dst; long di = idx[0] + X * (idx[1] + Y * (long)idx[2]);
          idx[d] = src; long si = idx[0] + X * (idx[1] + Y * (long)idx[2]);
          v[di] = v[si];
        }
      }
  }
}

void oracle_mask_bits(const nebo_mask_t* m, const int32_t* pts, int n, uint32_t* words)
{
  long npts = (long)m->glob[0] * m->glob[1] * m->glob[2];
  memset(words, 0, sizeof(uint32_t) * (size_t)((npts + 31) / 32));
  for (int i = 0; i < n; ++i) {
    long gx = m->offset[0] + m->ghost_minus[0] + pts[3 * i];
    long gy = m->offset[1] + m->ghost_minus[1] + pts[3 * i + 1];
    long gz = m->offset[2] + m->ghost_minus[2] + pts[3 * i + 2];
    long pos = gx + (long)m->glob[0] * (gy + (long)m->glob[1] * gz);
    words[pos / 32] |= 1u << (pos % 32);
  }
}

/* ------------------------------------------------------------------------- */
/* closed forms                                                              */
/* ------------------------------------------------------------------------- */

#define AT(f, x, y, z) (((const double*)(f)->values)[field_index((f), (x), (y), (z))])

/* face value F on the low face of cell (x,y,z) along axis d:
 * interp(k)*grad(phi) = (k[-1]*i0 + k[0]*i1) * (phi[-1]*g0 + phi[0]*g1) */
static inline double face_flux(const nebo_field_t* k, const nebo_field_t* phi,
                               const nebo_axis_coefs_t* c, int d, int x, int y, int z)
{
  int xm = x - (d == 0), ym = y - (d == 1), zm = z - (d == 2);
  double ki = AT(k, xm, ym, zm) * c->interp[d][0] + AT(k, x, y, z) * c->interp[d][1];
  double gr = AT(phi, xm, ym, zm) * c->grad[d][0] + AT(phi, x, y, z) * c->grad[d][1];
  return ki * gr;
}

static void divkgrad_range(const nebo_field_t* r, const nebo_field_t* k,
                           const nebo_field_t* phi, const nebo_axis_coefs_t* c, int form,
                           const nebo_limits_t* L)
{
  double* out = (double*)r->values;
  for (int z = L->lo[2]; z < L->hi[2]; ++z)
    for (int y = L->lo[1]; y < L->hi[1]; ++y)
      for (int x = L->lo[0]; x < L->hi[0]; ++x) {
        double D[3];
        for (int d = 0; d < 3; ++d) {
          double lo = face_flux(k, phi, c, d, x, y, z);
          double hi = face_flux(k, phi, c, d, x + (d == 0), y + (d == 1), z + (d == 2));
          D[d] = lo * c->div[d][0] + hi * c->div[d][1];
        }
        double v = form == 0 ? (D[0] + D[1]) + D[2] : ((-D[0]) - D[1]) - D[2];
        out[field_index(r, x, y, z)] = v;
      }
}

static void interior_limits(const nebo_field_t* f, nebo_limits_t* L)
{
  for (int d = 0; d < 3; ++d) {
    L->lo[d] = 0;
    L->hi[d] = f->extent[d] - f->valid_minus[d] - f->valid_plus[d];
  }
}

void oracle_divkgrad(const nebo_field_t* r, const nebo_field_t* k, const nebo_field_t* phi,
                     const nebo_axis_coefs_t* c, int form)
{
  nebo_limits_t L;
  interior_limits(r, &L);
  divkgrad_range(r, k, phi, c, form, &L);
}

static void* task_divkgrad(void* p)
{
  task_arg_t* t = (task_arg_t*)p;
  divkgrad_range(t->lhs, t->k, t->phi, t->c, t->form, &t->lim);
  return NULL;
}

void oracle_divkgrad_threads(const nebo_field_t* r, const nebo_field_t* k,
                             const nebo_field_t* phi, const nebo_axis_coefs_t* c, int form,
                             int nthreads)
{
  nebo_limits_t L;
  interior_limits(r, &L);
  task_arg_t* t;
  int n = make_tasks(&L, nthreads, &t);
  for (int i = 0; i < n; ++i) { t[i].lhs = r; t[i].k = k; t[i].phi = phi; t[i].c = c; t[i].form = form; }
  run_tasks(t, n, task_divkgrad);
  free(t);
}

void oracle_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                       const nebo_axis_coefs_t* c, double s)
{
  nebo_limits_t L;
  interior_limits(out, &L);
  double* o = (double*)out->values;
  for (int z = L.lo[2]; z < L.hi[2]; ++z)
    for (int y = L.lo[1]; y < L.hi[1]; ++y)
      for (int x = L.lo[0]; x < L.hi[0]; ++x) {
        double D[3];
        for (int d = 0; d < 3; ++d) {
          int dx = d == 0, dy = d == 1, dz = d == 2;
          double glo = AT(phi, x - dx, y - dy, z - dz) * c->grad[d][0] + AT(phi, x, y, z) * c->grad[d][1];
          double ghi = AT(phi, x, y, z) * c->grad[d][0] + AT(phi, x + dx, y + dy, z + dz) * c->grad[d][1];
          D[d] = glo * c->div[d][0] + ghi * c->div[d][1];
        }
        o[field_index(out, x, y, z)] = AT(phi, x, y, z) + s * ((D[0] + D[1]) + D[2]);
      }
}

/* ------------------------------------------------------------------------- */
/* adjoint of the periodic diffusion step, from separate operator passes      */
/* ------------------------------------------------------------------------- */

void oracle_diffusion_step(const nebo_field_t* out, const nebo_field_t* phi,
                           const nebo_field_t* k, const nebo_axis_coefs_t* c, double dt)
{
  /* r = div(interp(k) * grad(phi)) through the closed form, then Euler */
  nebo_limits_t L;
  interior_limits(out, &L);
  double* o = (double*)out->values;
  for (int z = L.lo[2]; z < L.hi[2]; ++z)
    for (int y = L.lo[1]; y < L.hi[1]; ++y)
      for (int x = L.lo[0]; x < L.hi[0]; ++x) {
        double D[3];
        for (int d = 0; d < 3; ++d) {
          double lo = face_flux(k, phi, c, d, x, y, z);
          double hi = face_flux(k, phi, c, d, x + (d == 0), y + (d == 1), z + (d == 2));
          D[d] = lo * c->div[d][0] + hi * c->div[d][1];
        }
        o[field_index(out, x, y, z)] = AT(phi, x, y, z) + dt * ((D[0] + D[1]) + D[2]);
      }
  oracle_fill_ghosts_periodic(out);
}

void oracle_diffusion_adjoint_step(const nebo_field_t* lam_out, double* gk,
                                   const nebo_field_t* lam, const nebo_field_t* phi,
                                   const nebo_field_t* k, const nebo_axis_coefs_t* c,
                                   double dt)
{
  /* Separate passes over face temporaries, one per axis:
   *   w  = D^T lam     on faces: w_i = lam_i*d0 + lam_{i-1}*d1
   *   gp = G phi       on faces: gp_i = phi_{i-1}*g0 + phi_i*g1
   *   kf = I k         on faces: kf_i = k_{i-1}*i0 + k_i*i1
   *   a  = G^T (kf*w)  on cells: a_c = (kf*w)_{c+1}*g0 + (kf*w)_c*g1
   *   b  = I^T (w*gp)  on cells: b_c = (w*gp)_{c+1}*i0 + (w*gp)_c*i1
   * lam_out = lam + dt*((ax+ay)+az);  gk += dt*((bx+by)+bz).
   * Face i is the low face of cell i; the n interior cells use faces 0..n. */
  nebo_limits_t L;
  interior_limits(lam_out, &L);
  int nx = L.hi[0], ny = L.hi[1], nz = L.hi[2];
  size_t ncell = (size_t)nx * ny * nz;
  double* A[3];
  double* B[3];
  for (int d = 0; d < 3; ++d) {
    int fx = nx + (d == 0), fy = ny + (d == 1), fz = nz + (d == 2);
    size_t nf = (size_t)fx * fy * fz;
    double* w = (double*)malloc(nf * sizeof(double));
    double* kw = (double*)malloc(nf * sizeof(double));
    double* wg = (double*)malloc(nf * sizeof(double));
    int dx = d == 0, dy = d == 1, dz = d == 2;
    for (int z = 0; z < fz; ++z)
      for (int y = 0; y < fy; ++y)
        for (int x = 0; x < fx; ++x) {
          size_t fi = (size_t)x + (size_t)fx * (y + (size_t)fy * z);
          w[fi] = AT(lam, x, y, z) * c->div[d][0] + AT(lam, x - dx, y - dy, z - dz) * c->div[d][1];
          double gp = AT(phi, x - dx, y - dy, z - dz) * c->grad[d][0] + AT(phi, x, y, z) * c->grad[d][1];
          double kf = AT(k, x - dx, y - dy, z - dz) * c->interp[d][0] + AT(k, x, y, z) * c->interp[d][1];
          kw[fi] = kf * w[fi];
          wg[fi] = w[fi] * gp;
        }
    A[d] = (double*)malloc(ncell * sizeof(double));
    B[d] = (double*)malloc(ncell * sizeof(double));
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          size_t ci = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
          size_t flo = (size_t)x + (size_t)fx * (y + (size_t)fy * z);
          size_t fhi = (size_t)(x + dx) + (size_t)fx * ((y + dy) + (size_t)fy * (z + dz));
          A[d][ci] = kw[fhi] * c->grad[d][0] + kw[flo] * c->grad[d][1];
          B[d][ci] = wg[fhi] * c->interp[d][0] + wg[flo] * c->interp[d][1];
        }
    free(w); free(kw); free(wg);
  }
  double* lo = (double*)lam_out->values;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t ci = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
        size_t fi = field_index(lam_out, x, y, z);
        lo[fi] = AT(lam, x, y, z) + dt * ((A[0][ci] + A[1][ci]) + A[2][ci]);
        size_t gi = field_index(k, x, y, z);
        gk[gi] = gk[gi] + dt * ((B[0][ci] + B[1][ci]) + B[2][ci]);
      }
  for (int d = 0; d < 3; ++d) { free(A[d]); free(B[d]); }
  oracle_fill_ghosts_periodic(lam_out);
}

/* ------------------------------------------------------------------------- */
/* deterministic inputs                                                      */
/* ------------------------------------------------------------------------- */

uint64_t oracle_splitmix64(uint64_t* s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void oracle_fill_uniform(double* dst, size_t n, uint64_t seed, double lo, double hi)
{
  uint64_t s = seed;
  for (size_t i = 0; i < n; ++i) {
    double u = (double)(oracle_splitmix64(&s) >> 11) * 0x1.0p-53;
    dst[i] = lo + (hi - lo) * u;
  }
}

void oracle_fill_normal(double* dst, size_t n, uint64_t seed)
{
  uint64_t s = seed;
  for (size_t i = 0; i < n; i += 2) {
    double u1 = ((double)(oracle_splitmix64(&s) >> 11) + 1.0) * 0x1.0p-53;
    double u2 = (double)(oracle_splitmix64(&s) >> 11) * 0x1.0p-53;
    double r = sqrt(-2.0 * log(u1));
    dst[i] = r * cos(6.283185307179586 * u2);
    if (i + 1 < n) dst[i + 1] = r * sin(6.283185307179586 * u2);
  }
}

/* ------------------------------------------------------------------------- */
/* scalar transport: upwinded convection + var-coefficient diffusion, SSP-RK3 */
/* ------------------------------------------------------------------------- */

/* The right-hand side R(s) = div(interp(k)*grad(s)) - div(u*upwind(s)) at one
 * interior cell, in the order the fused expression
 *   (Dx(Ix(k)*Gx(s)) + Dy(..)) + Dz(..)  -  ((Dx(c_x*u_x) + Dy(..)) + Dz(..)),
 *   c_d = cond(u_d > 0, s[-1 along d])(s[0])
 * is evaluated by the tree walk (NeboStencils.h:407-448 fold, NeboCond.h:408-412
 * select, NeboOperators.h SumOp/DiffOp/ProdOp), where every one-point shift
 * stencil multiplies by its coefficient 1.0. */
static double transport_rhs_cell(const nebo_field_t* s, const nebo_field_t* k,
                                 const nebo_field_t* u, const nebo_axis_coefs_t* c,
                                 int x, int y, int z)
{
  double f[3], v[3];
  for (int d = 0; d < 3; ++d) {
    int dx = d == 0, dy = d == 1, dz = d == 2;
    double lo = face_flux(k, s, c, d, x, y, z);
    double hi = face_flux(k, s, c, d, x + dx, y + dy, z + dz);
    f[d] = lo * c->div[d][0] + hi * c->div[d][1];
    const nebo_field_t* ud = &u[d];
    double ul = AT(ud, x, y, z), uh = AT(ud, x + dx, y + dy, z + dz);
    double cl = ul > 0.0 ? AT(s, x - dx, y - dy, z - dz) * 1.0 : AT(s, x, y, z) * 1.0;
    double ch = uh > 0.0 ? AT(s, x, y, z) * 1.0 : AT(s, x + dx, y + dy, z + dz) * 1.0;
    v[d] = (cl * ul) * c->div[d][0] + (ch * uh) * c->div[d][1];
  }
  return ((f[0] + f[1]) + f[2]) - ((v[0] + v[1]) + v[2]);
}

static void transport_stage_range(const nebo_field_t* out, const nebo_field_t* phin,
                                  const nebo_field_t* s, const nebo_field_t* k, const nebo_field_t* u,
                                  const nebo_axis_coefs_t* c, double dt, int stage, const nebo_limits_t* L)
{
  double* o = (double*)out->values;
  for (int z = L->lo[2]; z < L->hi[2]; ++z)
    for (int y = L->lo[1]; y < L->hi[1]; ++y)
      for (int x = L->lo[0]; x < L->hi[0]; ++x) {
        double r = transport_rhs_cell(s, k, u, c, x, y, z);
        double pn = AT(phin, x, y, z), sv = AT(s, x, y, z);
        double v;
        if (stage == 1) v = pn + dt * r;
        else if (stage == 2) v = 0.75 * pn + 0.25 * (sv + dt * r);
        else v = (1.0 / 3.0) * pn + (2.0 / 3.0) * (sv + dt * r);
        o[field_index(out, x, y, z)] = v;
      }
}

void oracle_transport_stage(const nebo_field_t* out, const nebo_field_t* phin,
                            const nebo_field_t* s, const nebo_field_t* k, const nebo_field_t* u,
                            const nebo_axis_coefs_t* c, double dt, int stage)
{
  nebo_limits_t L;
  interior_limits(out, &L);
  transport_stage_range(out, phin, s, k, u, c, dt, stage, &L);
  oracle_fill_ghosts_periodic(out);
}

typedef struct transport_task {
  const nebo_field_t *out, *phin, *s, *k, *u;
  const nebo_axis_coefs_t* c;
  double dt;
  int stage;
  nebo_limits_t lim;
} transport_task_t;

static void* task_transport(void* p)
{
  transport_task_t* t = (transport_task_t*)p;
  transport_stage_range(t->out, t->phin, t->s, t->k, t->u, t->c, t->dt, t->stage, &t->lim);
  return NULL;
}

/* The same stage with the interior split into the z slabs of the thread-parallel
 * walk (NeboBasic.h:112-191); each cell is computed by the same code, so the
 * result equals oracle_transport_stage bit for bit. */
void oracle_transport_stage_threads(const nebo_field_t* out, const nebo_field_t* phin,
                                    const nebo_field_t* s, const nebo_field_t* k, const nebo_field_t* u,
                                    const nebo_axis_coefs_t* c, double dt, int stage, int nthreads)
{
  nebo_limits_t L;
  interior_limits(out, &L);
  task_arg_t* parts;
  int n = make_tasks(&L, nthreads, &parts);
  transport_task_t* t = (transport_task_t*)calloc((size_t)n, sizeof(transport_task_t));
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    transport_task_t v = {out, phin, s, k, u, c, dt, stage, parts[i].lim};
    t[i] = v;
  }
  for (int i = 1; i < n; ++i) pthread_create(&th[i], NULL, task_transport, &t[i]);
  task_transport(&t[0]);
  for (int i = 1; i < n; ++i) pthread_join(th[i], NULL);
  free(th);
  free(t);
  free(parts);
  oracle_fill_ghosts_periodic(out);
}

/* Transpose of one SSP-RK3 stage.  With R_s the (linear in s) right-hand side
 * for fixed k and u, and R_k(s) its (linear in k) diffusion part for fixed s:
 *   R_s^T a = sum_d G_d^T(kf_d .* w_d) - sum_d U_d^T(u_d .* w_d),  w_d = D_d^T a
 *   R_k(s)^T a = sum_d I_d^T(w_d .* G_d s)
 * with, for face f = low face of cell f along d,
 *   (D^T a)_f = a_f*d0 + a_{f-1}*d1        (G^T v)_c = v_{c+1}*g0 + v_c*g1
 *   (I^T v)_c = v_{c+1}*i0 + v_c*i1        (U^T v)_c = [u_{c+1} > 0] v_{c+1} + [!(u_c > 0)] v_c
 * and X = a + dt*R_s^T a, the stage outputs are
 *   stage 3: out = (2/3)*X,          acc_out = (1/3)*a,        gk += ((2/3)*dt) * R_k(s)^T a
 *   stage 2: out = 0.25*X,           acc_out = acc_in + 0.75*a, gk += (0.25*dt) * R_k(s)^T a
 *   stage 1: out = acc_in + X,                                  gk += dt * R_k(s)^T a
 * Separate passes over face temporaries, one per axis; the periodic ghosts of a,
 * s and k hold the wrapped interior, so faces 0 and n are one physical face. */
void oracle_transport_adjoint_stage(const nebo_field_t* out, double* acc_out, double* gk,
                                    const nebo_field_t* a, const double* acc_in,
                                    const nebo_field_t* s, const nebo_field_t* k,
                                    const nebo_field_t* u, const nebo_axis_coefs_t* c, double dt,
                                    int stage)
{
  nebo_limits_t L;
  interior_limits(out, &L);
  int nx = L.hi[0], ny = L.hi[1], nz = L.hi[2];
  size_t ncell = (size_t)nx * ny * nz;
  double* A[3];
  double* V[3];
  double* B[3];
  for (int d = 0; d < 3; ++d) {
    int fx = nx + (d == 0), fy = ny + (d == 1), fz = nz + (d == 2);
    size_t nf = (size_t)fx * fy * fz;
    double* kw = (double*)malloc(nf * sizeof(double));
    double* wg = (double*)malloc(nf * sizeof(double));
    double* uw = (double*)malloc(nf * sizeof(double));
    double* uf = (double*)malloc(nf * sizeof(double));
    int dx = d == 0, dy = d == 1, dz = d == 2;
    for (int z = 0; z < fz; ++z)
      for (int y = 0; y < fy; ++y)
        for (int x = 0; x < fx; ++x) {
          size_t fi = (size_t)x + (size_t)fx * (y + (size_t)fy * z);
          double w = AT(a, x, y, z) * c->div[d][0] + AT(a, x - dx, y - dy, z - dz) * c->div[d][1];
          double gs = AT(s, x - dx, y - dy, z - dz) * c->grad[d][0] + AT(s, x, y, z) * c->grad[d][1];
          double kf = AT(k, x - dx, y - dy, z - dz) * c->interp[d][0] + AT(k, x, y, z) * c->interp[d][1];
          uf[fi] = AT(&u[d], x, y, z);
          kw[fi] = kf * w;
          wg[fi] = w * gs;
          uw[fi] = uf[fi] * w;
        }
    A[d] = (double*)malloc(ncell * sizeof(double));
    V[d] = (double*)malloc(ncell * sizeof(double));
    B[d] = (double*)malloc(ncell * sizeof(double));
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          size_t ci = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
          size_t flo = (size_t)x + (size_t)fx * (y + (size_t)fy * z);
          size_t fhi = (size_t)(x + dx) + (size_t)fx * ((y + dy) + (size_t)fy * (z + dz));
          A[d][ci] = kw[fhi] * c->grad[d][0] + kw[flo] * c->grad[d][1];
          V[d][ci] = (uf[fhi] > 0.0 ? uw[fhi] : 0.0) + (uf[flo] > 0.0 ? 0.0 : uw[flo]);
          B[d][ci] = wg[fhi] * c->interp[d][0] + wg[flo] * c->interp[d][1];
        }
    free(kw); free(wg); free(uw); free(uf);
  }
  const double cs = stage == 3 ? 2.0 / 3.0 : stage == 2 ? 0.25 : 1.0;
  const double cdt = cs * dt;
  double* o = (double*)out->values;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        size_t ci = (size_t)x + (size_t)nx * (y + (size_t)ny * z);
        size_t fi = field_index(out, x, y, z);
        double av = AT(a, x, y, z);
        double rt = ((A[0][ci] + A[1][ci]) + A[2][ci]) - ((V[0][ci] + V[1][ci]) + V[2][ci]);
        double X = av + dt * rt;
        if (stage == 3) {
          o[fi] = cs * X;
          acc_out[fi] = (1.0 / 3.0) * av;
        } else if (stage == 2) {
          o[fi] = cs * X;
          acc_out[fi] = acc_in[fi] + 0.75 * av;
        } else {
          o[fi] = acc_in[fi] + X;
        }
        size_t gi = field_index(k, x, y, z);
        gk[gi] = gk[gi] + cdt * ((B[0][ci] + B[1][ci]) + B[2][ci]);
      }
  for (int d = 0; d < 3; ++d) { free(A[d]); free(V[d]); free(B[d]); }
  oracle_fill_ghosts_periodic(out);
}
