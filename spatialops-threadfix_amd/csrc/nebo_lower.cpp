// nebo_lower.cpp -- host half of the engine: ghost inference, expression
// validation, lowering of an expression tree to the device stack program, and
// recognition of the trees that have hand-written kernels.
//
// Ghost rules follow spatialops/structured/GhostData.h:54-257 (GHOST_MAX is
// "infinite": an infinite left operand of + stays, an infinite right operand
// wins), leaf ghosts of spatialops/NeboRhs.h:52-54,183-188, stencil ghosts of
// spatialops/NeboStencils.h:222-239,311-317 (each point p moves the minus count
// by p and the plus count by -p, folded with min), edgeless stencils of
// NeboStencils.h:263-276,556-567 (additive_point_to_ghost: a point only widens the
// side opposite to the one it reads), mask shifts of NeboStencils.h:1048-1051,
// cond of NeboCond.h:86-91,284-289,
// and the limits of calculate_actual_ghost (NeboBasic.h:61-100) and
// NeboField<Initial>::assign (NeboLhs.h:67-88).

#include "nebo_internal.h"

#include <cstring>
#include <map>
#include <sstream>
#include <tuple>

namespace nebo {

namespace {

Ghost g_add(Ghost a, const Ghost& b)
{
  if (a.infinite()) return a;
  if (b.infinite()) return b;
  for (int d = 0; d < 3; ++d) { a.m[d] += b.m[d]; a.p[d] += b.p[d]; }
  return a;
}

Ghost g_sub(Ghost a, const Ghost& b)
{
  for (int d = 0; d < 3; ++d) { a.m[d] -= b.m[d]; a.p[d] -= b.p[d]; }
  return a;
}

Ghost g_min(Ghost a, const Ghost& b)
{
  for (int d = 0; d < 3; ++d) {
    a.m[d] = b.m[d] < a.m[d] ? b.m[d] : a.m[d];
    a.p[d] = b.p[d] < a.p[d] ? b.p[d] : a.p[d];
  }
  return a;
}

Ghost g_point(const int32_t v[3])
{
  Ghost g;
  for (int d = 0; d < 3; ++d) {
    g.m[d] = v[d] < 0 ? -v[d] : 0;
    g.p[d] = v[d] > 0 ? v[d] : 0;
  }
  return g;
}

Ghost stencil_shift(const nebo_stencil_t& s)
{
  Ghost g = Ghost::uniform(0);
  for (int i = 0; i < s.npts; ++i)
    for (int d = 0; d < 3; ++d) {
      const int mn = s.pts[i][d], pl = -s.pts[i][d];
      if (i == 0 || mn < g.m[d]) g.m[d] = mn;
      if (i == 0 || pl < g.p[d]) g.p[d] = pl;
    }
  return g;
}

// Edgeless rule: additive_point_to_ghost(p) = (minus max(p, 0), plus max(-p, 0)),
// folded with min over the points.  The reference only compiles the one-point
// form (its multi-point fold at NeboStencils.h:231-233 calls a misspelt
// function); the fold here is the one its other point collections use.
Ghost edgeless_shift(const nebo_stencil_t& s)
{
  Ghost g = Ghost::uniform(0);
  for (int i = 0; i < s.npts; ++i)
    for (int d = 0; d < 3; ++d) {
      const int v = s.pts[i][d];
      const int mn = v > 0 ? v : 0, pl = v < 0 ? -v : 0;
      if (i == 0 || mn < g.m[d]) g.m[d] = mn;
      if (i == 0 || pl < g.p[d]) g.p[d] = pl;
    }
  return g;
}

bool is_unary(int op)
{
  switch (op) {
  case NEBO_NEG: case NEBO_ABS: case NEBO_SQRT: case NEBO_SQUARE: case NEBO_CUBE:
  case NEBO_POWINT: case NEBO_EXP: case NEBO_LOG: case NEBO_LOG10: case NEBO_SIN:
  case NEBO_COS: case NEBO_TAN: case NEBO_TANH: case NEBO_ERF: case NEBO_ERFC:
  case NEBO_INV_ERF: case NEBO_INV_ERFC: case NEBO_NOT: case NEBO_STENCIL: case NEBO_MASK_SHIFT:
    return true;
  default:
    return false;
  }
}

bool is_leaf(int op) { return op == NEBO_SCALAR || op == NEBO_FIELD || op == NEBO_MASK; }

}  // namespace

int validate_expr(const nebo_expr_t* e, const char* ctx)
{
  if (!e || e->n_nodes <= 0 || !e->nodes)
    return set_error(NEBO_ERR_INVALID, ctx, "empty expression");
  for (int i = 0; i < e->n_nodes; ++i) {
    const nebo_node_t& n = e->nodes[i];
    std::ostringstream msg;
    if (n.op < 0 || n.op >= NEBO_NUM_OPCODES) {
      msg << "node " << i << " has unknown opcode " << n.op;
      return set_error(NEBO_ERR_INVALID, ctx, msg.str());
    }
    auto child_ok = [&](int c) { return c >= 0 && c < i; };
    if (n.op == NEBO_FIELD && (n.idx < 0 || n.idx >= e->n_fields || !e->fields)) {
      msg << "node " << i << " names field " << n.idx << " of " << e->n_fields;
      return set_error(NEBO_ERR_INVALID, ctx, msg.str());
    }
    if (n.op == NEBO_MASK && (n.idx < 0 || n.idx >= e->n_masks || !e->masks)) {
      msg << "node " << i << " names mask " << n.idx << " of " << e->n_masks;
      return set_error(NEBO_ERR_INVALID, ctx, msg.str());
    }
    if (n.op == NEBO_STENCIL) {
      if (n.idx < 0 || n.idx >= e->n_stencils || !e->stencils) {
        msg << "node " << i << " names stencil " << n.idx << " of " << e->n_stencils;
        return set_error(NEBO_ERR_INVALID, ctx, msg.str());
      }
      const nebo_stencil_t& s = e->stencils[n.idx];
      if (s.npts < 1 || s.npts > NEBO_MAX_STENCIL_POINTS ||
          (s.kind != NEBO_STENCIL_WEIGHTED && s.kind != NEBO_STENCIL_SUM &&
           s.kind != NEBO_STENCIL_EDGELESS)) {
        msg << "stencil " << n.idx << " has " << s.npts << " points, kind " << s.kind;
        return set_error(NEBO_ERR_INVALID, ctx, msg.str());
      }
    }
    if (n.op == NEBO_MASK_SHIFT) {
      if (n.idx < 0 || n.idx >= e->n_stencils || !e->stencils || e->stencils[n.idx].npts != 1) {
        msg << "mask shift node " << i << " needs a one-point stencil entry (got index " << n.idx
            << ")";
        return set_error(NEBO_ERR_INVALID, ctx, msg.str());
      }
      if (!child_ok(n.a) || e->nodes[n.a].op != NEBO_MASK) {
        msg << "mask shift node " << i << " must take a mask node that precedes it";
        return set_error(NEBO_ERR_INVALID, ctx, msg.str());
      }
    }
    if (is_leaf(n.op)) continue;
    if (!child_ok(n.a) || (!is_unary(n.op) && !child_ok(n.b)) ||
        (n.op == NEBO_COND && !child_ok(n.c))) {
      msg << "node " << i << " has a child index that does not precede it";
      return set_error(NEBO_ERR_INVALID, ctx, msg.str());
    }
  }
  return NEBO_OK;
}

int node_ghosts(const nebo_expr_t* e, int node, bool with_bc, Ghost* out)
{
  const nebo_node_t& n = e->nodes[node];
  Ghost a, b, c;
  int rc;
  switch (n.op) {
  case NEBO_SCALAR:
    *out = Ghost::uniform(NEBO_GHOST_MAX);
    return NEBO_OK;
  case NEBO_FIELD: {
    const nebo_field_t& f = e->fields[n.idx];
    Ghost g{{f.valid_minus[0], f.valid_minus[1], f.valid_minus[2]},
            {f.valid_plus[0], f.valid_plus[1], f.valid_plus[2]}};
    *out = with_bc ? g_add(g, g_point(f.bc_extra)) : g;
    return NEBO_OK;
  }
  case NEBO_MASK: {
    const nebo_mask_t& m = e->masks[n.idx];
    Ghost g{{m.valid_minus[0], m.valid_minus[1], m.valid_minus[2]},
            {m.valid_plus[0], m.valid_plus[1], m.valid_plus[2]}};
    *out = with_bc ? g_add(g, g_point(m.bc_extra)) : g;
    return NEBO_OK;
  }
  case NEBO_STENCIL: {
    if ((rc = node_ghosts(e, n.a, with_bc, &a))) return rc;
    const nebo_stencil_t& s = e->stencils[n.idx];
    *out = g_add(a, s.kind == NEBO_STENCIL_EDGELESS ? edgeless_shift(s) : stencil_shift(s));
    return NEBO_OK;
  }
  case NEBO_MASK_SHIFT:   // additive_reductive_point_to_ghost(P) = (P, -P)
    if ((rc = node_ghosts(e, n.a, with_bc, &a))) return rc;
    *out = g_add(a, stencil_shift(e->stencils[n.idx]));
    return NEBO_OK;
  case NEBO_COND:
    if ((rc = node_ghosts(e, n.a, with_bc, &a))) return rc;
    if ((rc = node_ghosts(e, n.b, with_bc, &b))) return rc;
    if ((rc = node_ghosts(e, n.c, with_bc, &c))) return rc;
    *out = g_min(g_min(a, b), c);
    return NEBO_OK;
  default:
    if (is_unary(n.op)) return node_ghosts(e, n.a, with_bc, out);
    if ((rc = node_ghosts(e, n.a, with_bc, &a))) return rc;
    if ((rc = node_ghosts(e, n.b, with_bc, &b))) return rc;
    *out = g_min(a, b);
    return NEBO_OK;
  }
}

void collect_reads(const nebo_expr_t* e, int node, int sx, int sy, int sz, bool edgeless,
                   std::vector<FieldRead>* out)
{
  const nebo_node_t& n = e->nodes[node];
  switch (n.op) {
  case NEBO_SCALAR: case NEBO_MASK: case NEBO_MASK_SHIFT:
    return;
  case NEBO_FIELD: {
    for (const FieldRead& r : *out)
      if (r.field == n.idx && r.s[0] == sx && r.s[1] == sy && r.s[2] == sz && r.edgeless == edgeless)
        return;
    out->push_back(FieldRead{n.idx, {sx, sy, sz}, edgeless});
    return;
  }
  case NEBO_STENCIL: {
    const nebo_stencil_t& s = e->stencils[n.idx];
    const bool eg = edgeless || s.kind == NEBO_STENCIL_EDGELESS;
    for (int i = 0; i < s.npts; ++i)
      collect_reads(e, n.a, sx + s.pts[i][0], sy + s.pts[i][1], sz + s.pts[i][2], eg, out);
    return;
  }
  default:
    collect_reads(e, n.a, sx, sy, sz, edgeless, out);
    if (is_unary(n.op)) return;
    collect_reads(e, n.b, sx, sy, sz, edgeless, out);
    if (n.op == NEBO_COND) collect_reads(e, n.c, sx, sy, sz, edgeless, out);
    return;
  }
}

namespace {

// has_extents()/extents(): the first operand that has extents decides.
bool node_extents(const nebo_expr_t* e, int node, int ext[3])
{
  const nebo_node_t& n = e->nodes[node];
  switch (n.op) {
  case NEBO_SCALAR:
    return false;
  case NEBO_FIELD: {
    const nebo_field_t& f = e->fields[n.idx];
    for (int d = 0; d < 3; ++d) ext[d] = f.extent[d] - f.valid_minus[d] - f.valid_plus[d];
    return true;
  }
  case NEBO_MASK: {
    const nebo_mask_t& m = e->masks[n.idx];
    for (int d = 0; d < 3; ++d) ext[d] = m.extent[d] - m.ghost_minus[d] - m.ghost_plus[d];
    return true;
  }
  case NEBO_COND:
    return node_extents(e, n.a, ext) || node_extents(e, n.b, ext) || node_extents(e, n.c, ext);
  default:
    if (is_unary(n.op)) return node_extents(e, n.a, ext);
    return node_extents(e, n.a, ext) || node_extents(e, n.b, ext);
  }
}

}  // namespace

int assign_limits(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                  nebo_limits_t* lim, int* empty)
{
  static const char* ctx = "Nebo Ghost Checking";
  int rc = validate_expr(rhs, "Nebo Assignment");
  if (rc) return rc;
  Ghost r;
  if ((rc = node_ghosts(rhs, rhs->n_nodes - 1, true, &r))) return rc;
  static const char* axis[3] = {"X", "Y", "Z"};
  for (int d = 0; d < 3; ++d)
    if (lhs->has_bc[d] && r.p[d] < lhs->bc_extra[d])
      return set_error(NEBO_ERR_GHOST, ctx,
                       std::string("Not enough valid extra cells to validate all interior "
                                   "cells in the ") + axis[d] + " direction");
  const Ghost extra = g_point(lhs->bc_extra);
  Ghost g;
  if (use_ghost) {
    Ghost l{{lhs->ghost_minus[0], lhs->ghost_minus[1], lhs->ghost_minus[2]},
            {lhs->ghost_plus[0], lhs->ghost_plus[1], lhs->ghost_plus[2]}};
    g = g_sub(g_min(g_add(l, extra), r), extra);
  } else {
    Ghost l{{0, 0, 0}, {lhs->bc_extra[0], lhs->bc_extra[1], lhs->bc_extra[2]}};
    g = g_sub(l, extra);
  }
  bool vol = true;
  for (int d = 0; d < 3; ++d) {
    const int ext = lhs->extent[d] - lhs->valid_minus[d] - lhs->valid_plus[d];
    lim->lo[d] = -g.m[d];
    lim->hi[d] = ext + g.p[d];
    if (lim->hi[d] - lim->lo[d] <= 0) vol = false;
  }
  *empty = vol ? 0 : 1;
  return NEBO_OK;
}

int reduce_limits(const nebo_expr_t* e, int use_ghost, nebo_limits_t* lim, int* empty)
{
  int rc = validate_expr(e, "Nebo Reduction");
  if (rc) return rc;
  Ghost g = Ghost::uniform(0);
  if (use_ghost && (rc = node_ghosts(e, e->n_nodes - 1, false, &g))) return rc;
  int ext[3] = {0, 0, 0};
  node_extents(e, e->n_nodes - 1, ext);
  // A tree that reads no field and no mask has "infinite" ghosts and no extent:
  // the reference's fold (NeboReductions.h:58-90) would walk
  // [-GHOST_MAX, GHOST_MAX)^3, 5.8e12 cells of one constant.  Refused here.
  for (int d = 0; d < 3; ++d)
    if (use_ghost && (g.m[d] >= NEBO_GHOST_MAX || g.p[d] >= NEBO_GHOST_MAX))
      return set_error(NEBO_ERR_INVALID, "Nebo Reduction",
                       "the expression reads no field or mask, so its region with ghost cells is unbounded; "
                       "reduce over the interior or include a field");
  bool vol = true;
  for (int d = 0; d < 3; ++d) {
    lim->lo[d] = -g.m[d];
    lim->hi[d] = ext[d] + g.p[d];
    if (lim->hi[d] - lim->lo[d] <= 0) vol = false;
  }
  *empty = vol ? 0 : 1;
  return NEBO_OK;
}

// ---------------------------------------------------------------------------
// lowering
// ---------------------------------------------------------------------------
namespace {

struct Lowerer {
  const nebo_expr_t* e;
  VmProgram* p;
  int depth = 0;
  int max_depth = 0;
  int guard = 0;   // > 0 inside an edgeless stencil: field reads are bounds-checked
  std::map<std::tuple<int, int, int, int>, int> load_slot;   // (field, sx, sy, sz)
  std::map<std::tuple<int, int, int, int>, int> mask_slot;
  std::map<std::pair<int, int>, int> classes;               // (X, Y) -> class
  std::map<uint64_t, int> const_slot;
  const char* ctx = "Nebo Assignment";
  int err = NEBO_OK;

  int fail(const std::string& msg)
  {
    if (!err) err = set_error(NEBO_ERR_TOO_LARGE, ctx, msg);
    return err;
  }

  int cls(int X, int Y)
  {
    auto it = classes.find({X, Y});
    if (it != classes.end()) return it->second;
    const int c = (int)classes.size();
    if (c >= kVmMaxClasses) { fail("too many distinct field shapes in one expression"); return 0; }
    classes[{X, Y}] = c;
    p->cls_X[c] = X;
    p->cls_XY[c] = X * Y;
    p->n_classes = c + 1;
    return c;
  }

  int konst(double v)
  {
    uint64_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    auto it = const_slot.find(bits);
    if (it != const_slot.end()) return it->second;
    const int k = (int)const_slot.size();
    if (k >= kVmMaxConsts) { fail("too many distinct constants in one expression"); return 0; }
    const_slot[bits] = k;
    p->consts[k] = v;
    return k;
  }

  void emit(int op, int imm, int pops, int pushes)
  {
    if (p->n_insns >= kVmMaxInsns) { fail("expression expands to too many operations"); return; }
    if (depth < pops) { fail("internal: stack underflow"); return; }
    VmInsn& in = p->insns[p->n_insns++];
    in.code = (uint16_t)((op << 4) | depth);
    in.imm = (int16_t)imm;
    depth += pushes - pops;
    if (depth > max_depth) max_depth = depth;
    if (depth > kVmMaxDepth) fail("expression needs too deep an evaluation stack");
  }

  int load(int f, int sx, int sy, int sz)
  {
    auto key = std::make_tuple(f, sx, sy, sz);
    auto it = load_slot.find(key);
    if (it != load_slot.end()) return it->second;
    const int s = (int)load_slot.size();
    if (s >= kVmMaxLoads) { fail("expression reads too many (field, offset) pairs"); return 0; }
    load_slot[key] = s;
    const nebo_field_t& fd = e->fields[f];
    double* o = frame_origin(&fd);
    double* b = o + sx + (long)fd.glob[0] * (sy + (long)fd.glob[1] * sz);
    p->loads[s].base = b;
    p->loads[s].cls = cls(fd.glob[0], fd.glob[1]);
    // base - values and the block size both fit in int32: fields are checked to
    // hold fewer than 2^31 points and stencil shifts are small
    const long off = (long)(b - static_cast<double*>(fd.values));
    const long npts = (long)fd.glob[0] * fd.glob[1] * fd.glob[2];
    p->loads[s].glo = (int32_t)(-off);
    p->loads[s].ghi = (int32_t)(npts - off);
    return s;
  }

  int mask(int m, int sx, int sy, int sz)
  {
    auto key = std::make_tuple(m, sx, sy, sz);
    auto it = mask_slot.find(key);
    if (it != mask_slot.end()) return it->second;
    const int s = (int)mask_slot.size();
    if (s >= kVmMaxMasks) { fail("expression reads too many mask offsets"); return 0; }
    mask_slot[key] = s;
    const nebo_mask_t& md = e->masks[m];
    VmMask& vm = p->masks[s];
    vm.bits = md.bits;
    const int sh[3] = {sx, sy, sz};
    for (int d = 0; d < 3; ++d) vm.o[d] = md.offset[d] + md.ghost_minus[d] + sh[d];
    vm.X = md.glob[0];
    vm.Y = md.glob[1];
    return s;
  }

  static int vm_binary(int op)
  {
    switch (op) {
    case NEBO_ADD: return VM_ADD; case NEBO_SUB: return VM_SUB;
    case NEBO_MUL: return VM_MUL; case NEBO_DIV: return VM_DIV;
    case NEBO_MAX: return VM_MAX; case NEBO_MIN: return VM_MIN;
    case NEBO_POW: return VM_POW;
    case NEBO_EQ: return VM_EQ; case NEBO_NE: return VM_NE; case NEBO_LT: return VM_LT;
    case NEBO_LE: return VM_LE; case NEBO_GT: return VM_GT; case NEBO_GE: return VM_GE;
    case NEBO_AND: return VM_AND; case NEBO_OR: return VM_OR;
    default: return -1;
    }
  }

  static int vm_unary(int op)
  {
    switch (op) {
    case NEBO_NEG: return VM_NEG; case NEBO_ABS: return VM_ABS; case NEBO_SQRT: return VM_SQRT;
    case NEBO_SQUARE: return VM_SQUARE; case NEBO_CUBE: return VM_CUBE;
    case NEBO_EXP: return VM_EXP; case NEBO_LOG: return VM_LOG; case NEBO_LOG10: return VM_LOG10;
    case NEBO_SIN: return VM_SIN; case NEBO_COS: return VM_COS; case NEBO_TAN: return VM_TAN;
    case NEBO_TANH: return VM_TANH; case NEBO_ERF: return VM_ERF; case NEBO_ERFC: return VM_ERFC;
    case NEBO_INV_ERF: return VM_INV_ERF; case NEBO_INV_ERFC: return VM_INV_ERFC;
    case NEBO_NOT: return VM_NOT;
    default: return -1;
    }
  }

  // Post-order expansion of node at offset (sx,sy,sz): same operations in the
  // same order as the reference's recursive eval, so the result is bit-identical.
  void gen(int node, int sx, int sy, int sz)
  {
    if (err) return;
    const nebo_node_t& n = e->nodes[node];
    switch (n.op) {
    case NEBO_SCALAR: emit(VM_CONST, konst(n.value), 0, 1); return;
    case NEBO_FIELD: emit(guard ? VM_LOADG : VM_LOAD, load(n.idx, sx, sy, sz), 0, 1); return;
    case NEBO_MASK: emit(VM_MASK, mask(n.idx, sx, sy, sz), 0, 1); return;
    case NEBO_MASK_SHIFT: {
      const int32_t* P = e->stencils[n.idx].pts[0];
      emit(VM_MASK, mask(e->nodes[n.a].idx, sx + P[0], sy + P[1], sz + P[2]), 0, 1);
      return;
    }
    case NEBO_STENCIL: {
      const nebo_stencil_t& s = e->stencils[n.idx];
      const bool edgeless = s.kind == NEBO_STENCIL_EDGELESS;
      guard += edgeless;
      for (int i = 0; i < s.npts; ++i) {
        gen(n.a, sx + s.pts[i][0], sy + s.pts[i][1], sz + s.pts[i][2]);
        if (s.kind != NEBO_STENCIL_SUM) emit(VM_MULC, konst(s.coefs[i]), 1, 1);
        if (i > 0) emit(VM_ADD, 0, 2, 1);
      }
      guard -= edgeless;
      return;
    }
    case NEBO_COND:
      gen(n.a, sx, sy, sz);
      gen(n.b, sx, sy, sz);
      gen(n.c, sx, sy, sz);
      emit(VM_SELECT, 0, 3, 1);
      return;
    case NEBO_POWINT:
      gen(n.a, sx, sy, sz);
      if (n.idx == 2) emit(VM_SQUARE, 0, 1, 1);   // correctly rounded x*x
      else if (n.idx == 1) { /* pow(x, 1) == x */ }
      else emit(VM_POWC, konst((double)n.idx), 1, 1);
      return;
    default: {
      const int u = vm_unary(n.op);
      if (u >= 0) { gen(n.a, sx, sy, sz); emit(u, 0, 1, 1); return; }
      const int b = vm_binary(n.op);
      if (b < 0) { fail("internal: unhandled opcode"); return; }
      gen(n.a, sx, sy, sz);
      gen(n.b, sx, sy, sz);
      emit(b, 0, 2, 1);
      return;
    }
    }
  }
};

}  // namespace

int lower_program(const nebo_expr_t* e, VmProgram* prog, int* max_depth)
{
  std::memset(prog, 0, sizeof *prog);
  Lowerer L;
  L.e = e;
  L.p = prog;
  L.gen(e->n_nodes - 1, 0, 0, 0);
  if (L.err) return L.err;
  if (L.depth != 1) return set_error(NEBO_ERR_INVALID, "Nebo Assignment", "internal: stack not balanced");
  if (prog->n_classes == 0) {  // scalar-only tree: give class 0 a shape anyway
    prog->n_classes = 1;
    prog->cls_X[0] = 1;
    prog->cls_XY[0] = 1;
  }
  *max_depth = L.max_depth;
  return NEBO_OK;
}

// ---------------------------------------------------------------------------
// recognised trees
// ---------------------------------------------------------------------------
namespace {

// A 2-point weighted stencil along axis d with points {lo*e_d, (lo+1)*e_d}.
bool two_point(const nebo_expr_t* e, int node, int d, int lo, double c[2], int* arg)
{
  const nebo_node_t& n = e->nodes[node];
  if (n.op != NEBO_STENCIL) return false;
  const nebo_stencil_t& s = e->stencils[n.idx];
  if (s.kind != NEBO_STENCIL_WEIGHTED || s.npts != 2) return false;
  for (int i = 0; i < 2; ++i)
    for (int a = 0; a < 3; ++a)
      if (s.pts[i][a] != (a == d ? lo + i : 0)) return false;
  c[0] = s.coefs[0];
  c[1] = s.coefs[1];
  *arg = n.a;
  return true;
}

int field_leaf(const nebo_expr_t* e, int node)
{
  const nebo_node_t& n = e->nodes[node];
  return n.op == NEBO_FIELD ? n.idx : -1;
}

bool same_field(const nebo_expr_t* e, int a, int b)
{
  if (a == b) return true;
  const nebo_field_t& x = e->fields[a];
  const nebo_field_t& y = e->fields[b];
  return std::memcmp(&x, &y, sizeof x) == 0;
}

// D_d = div_d( interp_d(k) * grad_d(phi) )   (either product order)
bool match_axis_flux(const nebo_expr_t* e, int node, int d, int* kf, int* pf,
                     nebo_axis_coefs_t* c)
{
  int prod;
  if (!two_point(e, node, d, 0, c->div[d], &prod)) return false;
  const nebo_node_t& m = e->nodes[prod];
  if (m.op != NEBO_MUL) return false;
  int ia, ga;
  for (int order = 0; order < 2; ++order) {
    const int ni = order ? m.b : m.a, ng = order ? m.a : m.b;
    if (two_point(e, ni, d, -1, c->interp[d], &ia) && two_point(e, ng, d, -1, c->grad[d], &ga)) {
      *kf = field_leaf(e, ia);
      *pf = field_leaf(e, ga);
      if (*kf >= 0 && *pf >= 0) return true;
    }
  }
  return false;
}

// L_d = div_d( grad_d(phi) )
bool match_axis_lapl(const nebo_expr_t* e, int node, int d, int* pf, nebo_axis_coefs_t* c)
{
  int g, a;
  if (!two_point(e, node, d, 0, c->div[d], &g)) return false;
  if (!two_point(e, g, d, -1, c->grad[d], &a)) return false;
  *pf = field_leaf(e, a);
  return *pf >= 0;
}

}  // namespace

bool match_divkgrad(const nebo_expr_t* e, DivKGradMatch* m)
{
  const int root = e->n_nodes - 1;
  const nebo_node_t& r = e->nodes[root];
  int ax[3];
  if (r.op == NEBO_ADD) {
    const nebo_node_t& l = e->nodes[r.a];
    if (l.op != NEBO_ADD) return false;
    ax[0] = l.a; ax[1] = l.b; ax[2] = r.b;
    m->form = 0;
  } else if (r.op == NEBO_SUB) {
    const nebo_node_t& l = e->nodes[r.a];
    if (l.op != NEBO_SUB) return false;
    const nebo_node_t& n = e->nodes[l.a];
    if (n.op != NEBO_NEG) return false;
    ax[0] = n.a; ax[1] = l.b; ax[2] = r.b;
    m->form = 1;
  } else {
    return false;
  }
  int kf[3], pf[3];
  for (int d = 0; d < 3; ++d)
    if (!match_axis_flux(e, ax[d], d, &kf[d], &pf[d], &m->c)) return false;
  for (int d = 1; d < 3; ++d)
    if (!same_field(e, kf[0], kf[d]) || !same_field(e, pf[0], pf[d])) return false;
  m->k_field = kf[0];
  m->phi_field = pf[0];
  return true;
}

bool match_lapl_euler(const nebo_expr_t* e, LaplEulerMatch* m)
{
  const int root = e->n_nodes - 1;
  const nebo_node_t& r = e->nodes[root];
  if (r.op != NEBO_ADD) return false;
  const int centre = field_leaf(e, r.a);
  if (centre < 0) return false;
  const nebo_node_t& prod = e->nodes[r.b];
  if (prod.op != NEBO_MUL) return false;
  int sum;
  if (e->nodes[prod.a].op == NEBO_SCALAR) { m->s = e->nodes[prod.a].value; sum = prod.b; }
  else if (e->nodes[prod.b].op == NEBO_SCALAR) { m->s = e->nodes[prod.b].value; sum = prod.a; }
  else return false;
  const nebo_node_t& s2 = e->nodes[sum];
  if (s2.op != NEBO_ADD) return false;
  const nebo_node_t& s1 = e->nodes[s2.a];
  if (s1.op != NEBO_ADD) return false;
  const int ax[3] = {s1.a, s1.b, s2.b};
  int pf[3];
  for (int d = 0; d < 3; ++d)
    if (!match_axis_lapl(e, ax[d], d, &pf[d], &m->c)) return false;
  for (int d = 0; d < 3; ++d)
    if (!same_field(e, centre, pf[d])) return false;
  for (int d = 0; d < 3; ++d) m->c.interp[d][0] = m->c.interp[d][1] = 0.0;
  m->phi_field = centre;
  return true;
}

}  // namespace nebo

// ---------------------------------------------------------------------------
// host-only C entry points
// ---------------------------------------------------------------------------
extern "C" int nebo_assign_limits(const nebo_field_t* lhs, const nebo_expr_t* rhs,
                                  int use_ghost, nebo_limits_t* limits, int* empty)
{
  nebo::clear_error();
  if (!lhs || !limits || !empty)
    return nebo::set_error(NEBO_ERR_INVALID, "Nebo Assignment", "null argument");
  return nebo::assign_limits(lhs, rhs, use_ghost, limits, empty);
}

extern "C" int nebo_assign_plan(const nebo_field_t* lhs, const nebo_expr_t* rhs,
                                int use_ghost, int* kernel_id)
{
  nebo::clear_error();
  if (!lhs || !kernel_id)
    return nebo::set_error(NEBO_ERR_INVALID, "Nebo Assignment", "null argument");
  int rc = nebo::validate_expr(rhs, "Nebo Assignment");
  if (rc) return rc;
  nebo::DivKGradMatch dm;
  nebo::LaplEulerMatch lm;
  *kernel_id = nebo::K_GENERIC;
  (void)use_ghost;
  if (nebo::match_divkgrad(rhs, &dm)) {
    const nebo_field_t& k = rhs->fields[dm.k_field];
    const nebo_field_t& p = rhs->fields[dm.phi_field];
    if (std::memcmp(k.glob, lhs->glob, sizeof k.glob) == 0 &&
        std::memcmp(p.glob, lhs->glob, sizeof p.glob) == 0)
      *kernel_id = nebo::K_DIVKGRAD;
  } else if (nebo::match_lapl_euler(rhs, &lm)) {
    const nebo_field_t& p = rhs->fields[lm.phi_field];
    if (std::memcmp(p.glob, lhs->glob, sizeof p.glob) == 0 && p.values != lhs->values)
      *kernel_id = nebo::K_LAPL_EULER;
  }
  return NEBO_OK;
}

extern "C" int nebo_lower_stats(const nebo_expr_t* rhs, int32_t* n_insns, int32_t* max_depth,
                                int32_t* n_loads)
{
  nebo::clear_error();
  int rc = nebo::validate_expr(rhs, "Nebo Assignment");
  if (rc) return rc;
  nebo::VmProgram prog;
  int depth = 0;
  if ((rc = nebo::lower_program(rhs, &prog, &depth))) return rc;
  if (n_insns) *n_insns = prog.n_insns;
  if (max_depth) *max_depth = depth;
  if (n_loads) {
    int n = 0;
    for (int i = 0; i < nebo::kVmMaxLoads; ++i)
      if (prog.loads[i].base) ++n;
    *n_loads = n;
  }
  return NEBO_OK;
}
