// Nebo.h -- the expression language of the C++ front end: the reference's user
// API (spatialops/Nebo.h and the headers it pulls in: NeboAssignment.h,
// NeboOperators.h, NeboCond.h, NeboMask.h, NeboStencils.h, NeboReductions.h),
// lowered onto the engine's C-ABI instead of the reference's GPUWalk templates.
//
// An expression is a small immutable tree; `lhs <<= rhs` flattens it into the
// post-order node list of nebo_expr_t and calls nebo_assign on the LHS's active
// GPU, which runs the whole tree as one fused kernel.  The field type of every
// operand is checked at compile time, as the reference's NeboFieldCheck does.
#pragma once

#include <spatialops/structured/SpatialField.h>

#include <array>
#include <functional>
#include <type_traits>
#include <unordered_map>

namespace SpatialOps {

namespace nebo_detail {

struct Node;
typedef std::shared_ptr<const Node> NodeP;

struct StencilSpec {
  int kind = NEBO_STENCIL_WEIGHTED;
  std::vector<std::array<int, 3>> pts;
  std::vector<double> coefs;
};

struct Node {
  int op = NEBO_SCALAR;
  NodeP a, b, c;
  double value = 0.0;
  int idx = 0;
  const void* leaf_id = nullptr;                       // field storage / mask identity
  std::function<nebo_field_t()> field;                 // descriptor at evaluation time
  std::function<bool(short, nebo_mask_t*)> mask;       // descriptor on a device
  std::shared_ptr<const StencilSpec> stencil;
};

inline NodeP scalar(double v)
{
  auto n = std::make_shared<Node>();
  n->op = NEBO_SCALAR;
  n->value = v;
  return n;
}

inline NodeP op1(int op, NodeP a, int idx = 0)
{
  auto n = std::make_shared<Node>();
  n->op = op;
  n->a = std::move(a);
  n->idx = idx;
  return n;
}

inline NodeP op2(int op, NodeP a, NodeP b)
{
  auto n = std::make_shared<Node>();
  n->op = op;
  n->a = std::move(a);
  n->b = std::move(b);
  return n;
}

inline NodeP op3(int op, NodeP a, NodeP b, NodeP c)
{
  auto n = std::make_shared<Node>();
  n->op = op;
  n->a = std::move(a);
  n->b = std::move(b);
  n->c = std::move(c);
  return n;
}

template <typename L, typename T>
NodeP field_leaf(const SpatialField<L, T>& f)
{
  auto n = std::make_shared<Node>();
  n->op = NEBO_FIELD;
  n->leaf_id = f.storage_id();
  SpatialField<L, T> h(f);  // shallow handle: shares storage
  n->field = [h]() { return h.descriptor(); };
  return n;
}

inline NodeP stencil(std::shared_ptr<const StencilSpec> s, NodeP arg)
{
  auto n = std::make_shared<Node>();
  n->op = NEBO_STENCIL;
  n->a = std::move(arg);
  n->stencil = std::move(s);
  return n;
}

// mask(x + P) of a mask leaf (NeboMaskShift, NeboStencils.h:1048-1213)
inline NodeP mask_shift(NodeP mask, const std::array<int, 3>& P)
{
  auto s = std::make_shared<StencilSpec>();
  s->pts.push_back(P);
  s->coefs.push_back(1.0);
  auto n = std::make_shared<Node>();
  n->op = NEBO_MASK_SHIFT;
  n->a = std::move(mask);
  n->stencil = std::move(s);
  return n;
}

}  // namespace nebo_detail

// ---------------------------------------------------------------------------
// SpatialMask (SpatialMask.h:99-181, BitField.h:291-362): a point list and its
// bit field over the prototype's glob block (bit b of word b/32 = flat index b).
// ---------------------------------------------------------------------------
template <typename FieldT>
class SpatialMask {
 public:
  typedef FieldT field_type;

  SpatialMask(const FieldT& prototype, const std::vector<IntVec>& points)
      : window_(prototype.window_with_ghost()),
        bc_(prototype.boundary_info()),
        ghosts_(prototype.get_ghost_data()),
        valid_(prototype.get_valid_ghost_data()),
        st_(std::make_shared<State>())
  {
    set_points(points);
  }

  // SpatialMask(window, bc, ghosts, points) (SpatialMask.h:99-114)
  SpatialMask(const MemoryWindow& window, const BoundaryCellInfo& bc, const GhostData& ghosts,
              const std::vector<IntVec>& points)
      : window_(window),
        bc_(bc.limit_by_extent(window.extent())),
        ghosts_(ghosts.limit_by_extent(window.extent())),
        valid_(ghosts_),
        st_(std::make_shared<State>())
  {
    set_points(points);
  }

  // SpatialMask<FieldT>::build(prototype, points) (SpatialMask.h:173-181): the
  // window a FieldT would have next to `prototype`, with its valid ghosts.
  template <typename Proto>
  static SpatialMask build(const Proto& prototype, const std::vector<IntVec>& points)
  {
    return SpatialMask(create_new_memory_window<FieldT, Proto>(prototype),
                       create_new_boundary_cell_info<FieldT, Proto>(prototype),
                       prototype.get_valid_ghost_data(), points);
  }

 private:
  void set_points(const std::vector<IntVec>& points)
  {
    const IntVec g = window_.glob_dim();
    st_->words.assign((window_.glob_npts() + 31) / 32, 0u);
    for (const IntVec& p : points) {
      const long x = window_.offset()[0] + ghosts_.get_minus(0) + p[0];
      const long y = window_.offset()[1] + ghosts_.get_minus(1) + p[1];
      const long z = window_.offset()[2] + ghosts_.get_minus(2) + p[2];
      const long pos = x + (long)g[0] * (y + (long)g[1] * z);
      st_->words[pos / 32] |= 1u << (pos % 32);
      st_->points.push_back(p[0]);
      st_->points.push_back(p[1]);
      st_->points.push_back(p[2]);
    }
  }

 public:

  // BitField::add_consumer (BitField.h:466-481): a copy of the bits on device d
  void add_consumer(short d)
  {
    if (IS_CPU_INDEX(d) || st_->dev.count(d)) return;
    void* p = nullptr;
    nebo_check(nebo_alloc(d, st_->words.size() * sizeof(uint32_t), &p));
    nebo_check(nebo_memcpy(p, st_->words.data(), st_->words.size() * sizeof(uint32_t), 0, nullptr));
    st_->dev[d] = std::shared_ptr<void>(p, [d](void* q) { nebo_free(d, q); });
  }
  bool find_consumer(short d) const { return IS_CPU_INDEX(d) || st_->dev.count(d) != 0; }
  bool operator()(int x, int y, int z) const
  {
    const IntVec g = window_.glob_dim();
    const long pos = (window_.offset()[0] + ghosts_.get_minus(0) + x) +
                     (long)g[0] * ((window_.offset()[1] + ghosts_.get_minus(1) + y) +
                                   (long)g[1] * (window_.offset()[2] + ghosts_.get_minus(2) + z));
    return (st_->words[pos / 32] >> (pos % 32)) & 1u;
  }
  const std::vector<int32_t>& flat_points() const { return st_->points; }
  const MemoryWindow& window_with_ghost() const { return window_; }
  const GhostData& get_ghost_data() const { return ghosts_; }
  const GhostData& get_valid_ghost_data() const { return valid_; }
  const BoundaryCellInfo& boundary_info() const { return bc_; }

  nebo_detail::NodeP node() const
  {
    auto n = std::make_shared<nebo_detail::Node>();
    n->op = NEBO_MASK;
    n->leaf_id = st_.get();
    SpatialMask<FieldT> self(*this);
    n->mask = [self](short d, nebo_mask_t* m) { return self.descriptor(d, m); };
    return n;
  }
  bool descriptor(short d, nebo_mask_t* m) const
  {
    auto it = st_->dev.find(d);
    if (it == st_->dev.end()) return false;
    std::memset(m, 0, sizeof *m);
    m->bits = static_cast<const uint32_t*>(it->second.get());
    for (int a = 0; a < 3; ++a) {
      m->glob[a] = window_.glob_dim()[a];
      m->offset[a] = window_.offset()[a];
      m->extent[a] = window_.extent()[a];
      m->ghost_minus[a] = ghosts_.get_minus(a);
      m->ghost_plus[a] = ghosts_.get_plus(a);
      m->valid_minus[a] = valid_.get_minus(a);
      m->valid_plus[a] = valid_.get_plus(a);
      m->has_bc[a] = bc_.has_bc(a);
      m->bc_extra[a] = bc_.has_extra(a);
    }
    return true;
  }

 private:
  struct State {
    std::vector<uint32_t> words;
    std::vector<int32_t> points;
    std::map<short, std::shared_ptr<void>> dev;
  };
  MemoryWindow window_;
  BoundaryCellInfo bc_;
  GhostData ghosts_, valid_;
  std::shared_ptr<State> st_;
};

namespace nebo_detail {

// Flattens a tree into nebo_expr_t (children before parents, root last).
class Flat {
 public:
  Flat(const NodeP& root, short device) : device_(device) { visit(root); }
  nebo_expr_t expr() const
  {
    nebo_expr_t e;
    e.n_nodes = (int32_t)nodes_.size();
    e.nodes = nodes_.data();
    e.n_fields = (int32_t)fields_.size();
    e.fields = fields_.data();
    e.n_stencils = (int32_t)stencils_.size();
    e.stencils = stencils_.data();
    e.n_masks = (int32_t)masks_.size();
    e.masks = masks_.data();
    return e;
  }

 private:
  int visit(const NodeP& n)
  {
    auto it = memo_.find(n.get());
    if (it != memo_.end()) return it->second;
    nebo_node_t out;
    std::memset(&out, 0, sizeof out);
    out.op = n->op;
    out.a = n->a ? visit(n->a) : -1;
    out.b = n->b ? visit(n->b) : -1;
    out.c = n->c ? visit(n->c) : -1;
    out.value = n->value;
    out.idx = n->idx;
    if (n->op == NEBO_FIELD) out.idx = table(field_ids_, n->leaf_id, [&] { fields_.push_back(n->field()); });
    if (n->op == NEBO_MASK)
      out.idx = table(mask_ids_, n->leaf_id, [&] {
        nebo_mask_t m;
        if (!n->mask(device_, &m))
          throw std::runtime_error(
              "Nebo error in Nebo Mask:\nmask used on a GPU without a copy there (call add_consumer)");
        masks_.push_back(m);
      });
    if (n->op == NEBO_STENCIL || n->op == NEBO_MASK_SHIFT)
      out.idx = table(stencil_ids_, n->stencil.get(), [&] {
        nebo_stencil_t s;
        std::memset(&s, 0, sizeof s);
        s.kind = n->stencil->kind;
        s.npts = (int32_t)n->stencil->pts.size();
        for (int i = 0; i < s.npts; ++i) {
          for (int d = 0; d < 3; ++d) s.pts[i][d] = n->stencil->pts[i][d];
          s.coefs[i] = i < (int)n->stencil->coefs.size() ? n->stencil->coefs[i] : 1.0;
        }
        stencils_.push_back(s);
      });
    nodes_.push_back(out);
    return memo_[n.get()] = (int)nodes_.size() - 1;
  }
  template <class F>
  int table(std::unordered_map<const void*, int>& ids, const void* key, F add)
  {
    auto it = ids.find(key);
    if (it != ids.end()) return it->second;
    const int i = (int)ids.size();
    add();
    ids[key] = i;
    return i;
  }

  short device_;
  std::vector<nebo_node_t> nodes_;
  std::vector<nebo_field_t> fields_;
  std::vector<nebo_stencil_t> stencils_;
  std::vector<nebo_mask_t> masks_;
  std::unordered_map<const void*, int> memo_, field_ids_, mask_ids_, stencil_ids_;
};

}  // namespace nebo_detail

// ---------------------------------------------------------------------------
// expression handles
// ---------------------------------------------------------------------------
namespace nebo_detail {
// What an expression handle carries: the lowered tree of the expression.  The
// reference's handles carry an expression-template object whose type encodes
// the tree; here the tree is data, so every expression has this one operand
// type and is evaluated by the engine behind the C-ABI.
class Lowered {
 public:
  explicit Lowered(NodeP n) : node_(std::move(n)) {}
  const NodeP& node() const { return node_; }

 private:
  NodeP node_;
};
}  // namespace nebo_detail

// NeboExpression<Operand, FieldType> / NeboBooleanExpression<Operand, FieldType>
// with the reference's typedefs and expr() (NeboBasic.h:193-243).
template <typename Operand, typename FieldType>
class NeboExpression {
 public:
  typedef FieldType field_type;
  typedef Operand Expression;
  explicit NeboExpression(const Operand& given) : expr_(given) {}
  template <typename O = Operand,
            typename = typename std::enable_if<std::is_same<O, nebo_detail::Lowered>::value>::type>
  explicit NeboExpression(nebo_detail::NodeP n) : expr_(std::move(n)) {}
  const Operand& expr() const { return expr_; }
  const nebo_detail::NodeP& node() const { return expr_.node(); }

 private:
  Operand expr_;
};

template <typename Operand, typename FieldType>
class NeboBooleanExpression {
 public:
  typedef FieldType field_type;
  typedef Operand Expression;
  explicit NeboBooleanExpression(const Operand& given) : expr_(given) {}
  template <typename O = Operand,
            typename = typename std::enable_if<std::is_same<O, nebo_detail::Lowered>::value>::type>
  explicit NeboBooleanExpression(nebo_detail::NodeP n) : expr_(std::move(n)) {}
  const Operand& expr() const { return expr_; }
  const nebo_detail::NodeP& node() const { return expr_.node(); }

 private:
  Operand expr_;
};

namespace nebo_detail {
template <typename FieldT> using Expr = NeboExpression<Lowered, FieldT>;
template <typename FieldT> using BoolExpr = NeboBooleanExpression<Lowered, FieldT>;
}  // namespace nebo_detail

namespace nebo_detail {

// operand traits: which field type an operand carries (void for scalars)
template <typename X, typename = void> struct Operand { static constexpr bool ok = false; };
template <typename X>
struct Operand<X, typename std::enable_if<std::is_arithmetic<X>::value>::type> {
  static constexpr bool ok = true, boolean = false;
  typedef void field;
  static NodeP node(X v) { return scalar((double)v); }
};
template <typename L, typename T>
struct Operand<SpatialField<L, T>> {
  static constexpr bool ok = true, boolean = false;
  typedef SpatialField<L, T> field;
  static NodeP node(const SpatialField<L, T>& f) { return field_leaf(f); }
};
template <typename R, typename F>
struct Operand<NeboExpression<R, F>> {
  static constexpr bool ok = true, boolean = false;
  typedef F field;
  static NodeP node(const NeboExpression<R, F>& e) { return e.node(); }
};
template <typename R, typename F>
struct Operand<NeboBooleanExpression<R, F>> {
  static constexpr bool ok = true, boolean = true;
  typedef F field;
  static NodeP node(const NeboBooleanExpression<R, F>& e) { return e.node(); }
};
template <typename F>
struct Operand<SpatialMask<F>> {
  static constexpr bool ok = true, boolean = true;
  typedef F field;
  static NodeP node(const SpatialMask<F>& m) { return m.node(); }
};

template <typename A, typename B>
struct Common {
  typedef typename Operand<A>::field FA;
  typedef typename Operand<B>::field FB;
  static_assert(std::is_void<FA>::value || std::is_void<FB>::value || std::is_same<FA, FB>::value,
                "Nebo: operands of one expression must have the same field type");
  typedef typename std::conditional<std::is_void<FA>::value, FB, FA>::type type;
};

template <typename A, typename B>
using ArithOK = typename std::enable_if<
    Operand<A>::ok && Operand<B>::ok && !Operand<A>::boolean && !Operand<B>::boolean &&
        !(std::is_arithmetic<A>::value && std::is_arithmetic<B>::value),
    nebo_detail::Expr<typename Common<A, B>::type>>::type;

template <typename A, typename B>
using CmpOK = typename std::enable_if<
    Operand<A>::ok && Operand<B>::ok && !Operand<A>::boolean && !Operand<B>::boolean &&
        !(std::is_arithmetic<A>::value && std::is_arithmetic<B>::value),
    nebo_detail::BoolExpr<typename Common<A, B>::type>>::type;

template <typename A, typename B>
using LogicOK = typename std::enable_if<Operand<A>::ok && Operand<B>::ok && Operand<A>::boolean &&
                                            Operand<B>::boolean,
                                        nebo_detail::BoolExpr<typename Common<A, B>::type>>::type;

template <typename A>
using UnaryOK = typename std::enable_if<Operand<A>::ok && !Operand<A>::boolean && !std::is_arithmetic<A>::value,
                                        nebo_detail::Expr<typename Operand<A>::field>>::type;

}  // namespace nebo_detail

#define NEBO_BINARY(OPSYM, CODE)                                                               \
  template <typename A, typename B>                                                            \
  nebo_detail::ArithOK<A, B> operator OPSYM(const A& a, const B& b)                            \
  {                                                                                            \
    return nebo_detail::ArithOK<A, B>(                                                         \
        nebo_detail::op2(CODE, nebo_detail::Operand<A>::node(a), nebo_detail::Operand<B>::node(b))); \
  }
NEBO_BINARY(+, NEBO_ADD)
NEBO_BINARY(-, NEBO_SUB)
NEBO_BINARY(*, NEBO_MUL)
NEBO_BINARY(/, NEBO_DIV)
#undef NEBO_BINARY

#define NEBO_COMPARE(OPSYM, CODE)                                                              \
  template <typename A, typename B>                                                            \
  nebo_detail::CmpOK<A, B> operator OPSYM(const A& a, const B& b)                              \
  {                                                                                            \
    return nebo_detail::CmpOK<A, B>(                                                           \
        nebo_detail::op2(CODE, nebo_detail::Operand<A>::node(a), nebo_detail::Operand<B>::node(b))); \
  }
NEBO_COMPARE(==, NEBO_EQ)
NEBO_COMPARE(!=, NEBO_NE)
NEBO_COMPARE(<, NEBO_LT)
NEBO_COMPARE(<=, NEBO_LE)
NEBO_COMPARE(>, NEBO_GT)
NEBO_COMPARE(>=, NEBO_GE)
#undef NEBO_COMPARE

template <typename A, typename B>
nebo_detail::LogicOK<A, B> operator&&(const A& a, const B& b)
{
  return nebo_detail::LogicOK<A, B>(
      nebo_detail::op2(NEBO_AND, nebo_detail::Operand<A>::node(a), nebo_detail::Operand<B>::node(b)));
}
template <typename A, typename B>
nebo_detail::LogicOK<A, B> operator||(const A& a, const B& b)
{
  return nebo_detail::LogicOK<A, B>(
      nebo_detail::op2(NEBO_OR, nebo_detail::Operand<A>::node(a), nebo_detail::Operand<B>::node(b)));
}
template <typename F>
nebo_detail::BoolExpr<F> operator!(const nebo_detail::BoolExpr<F>& a)
{
  return nebo_detail::BoolExpr<F>(nebo_detail::op1(NEBO_NOT, a.node()));
}
template <typename F>
nebo_detail::BoolExpr<F> operator!(const SpatialMask<F>& m)
{
  return nebo_detail::BoolExpr<F>(nebo_detail::op1(NEBO_NOT, m.node()));
}

template <typename A>
nebo_detail::UnaryOK<A> operator-(const A& a)
{
  return nebo_detail::UnaryOK<A>(nebo_detail::op1(NEBO_NEG, nebo_detail::Operand<A>::node(a)));
}

#define NEBO_FUNCTION(NAME, CODE)                                                               \
  template <typename A>                                                                         \
  nebo_detail::UnaryOK<A> NAME(const A& a)                                                      \
  {                                                                                             \
    return nebo_detail::UnaryOK<A>(nebo_detail::op1(CODE, nebo_detail::Operand<A>::node(a)));   \
  }
NEBO_FUNCTION(sin, NEBO_SIN)
NEBO_FUNCTION(cos, NEBO_COS)
NEBO_FUNCTION(tan, NEBO_TAN)
NEBO_FUNCTION(exp, NEBO_EXP)
NEBO_FUNCTION(tanh, NEBO_TANH)
NEBO_FUNCTION(abs, NEBO_ABS)
NEBO_FUNCTION(sqrt, NEBO_SQRT)
NEBO_FUNCTION(log, NEBO_LOG)
NEBO_FUNCTION(log10, NEBO_LOG10)
NEBO_FUNCTION(erf, NEBO_ERF)
NEBO_FUNCTION(erfc, NEBO_ERFC)
NEBO_FUNCTION(inv_erf, NEBO_INV_ERF)
NEBO_FUNCTION(inv_erfc, NEBO_INV_ERFC)
NEBO_FUNCTION(square, NEBO_SQUARE)
NEBO_FUNCTION(cube, NEBO_CUBE)
#undef NEBO_FUNCTION

template <typename A>
nebo_detail::UnaryOK<A> pow(const A& a, int n)
{
  return nebo_detail::UnaryOK<A>(nebo_detail::op1(NEBO_POWINT, nebo_detail::Operand<A>::node(a), n));
}
template <typename A, typename B>
typename std::enable_if<!std::is_integral<B>::value, nebo_detail::ArithOK<A, B>>::type pow(const A& a, const B& b)
{
  return nebo_detail::ArithOK<A, B>(
      nebo_detail::op2(NEBO_POW, nebo_detail::Operand<A>::node(a), nebo_detail::Operand<B>::node(b)));
}
template <typename A, typename B>
nebo_detail::ArithOK<A, B> max(const A& a, const B& b)
{
  return nebo_detail::ArithOK<A, B>(
      nebo_detail::op2(NEBO_MAX, nebo_detail::Operand<A>::node(a), nebo_detail::Operand<B>::node(b)));
}
template <typename A, typename B>
nebo_detail::ArithOK<A, B> min(const A& a, const B& b)
{
  return nebo_detail::ArithOK<A, B>(
      nebo_detail::op2(NEBO_MIN, nebo_detail::Operand<A>::node(a), nebo_detail::Operand<B>::node(b)));
}

// ---------------------------------------------------------------------------
// cond(test, value)(test, value)...(default)   (NeboCond.h: first true clause)
// ---------------------------------------------------------------------------
template <typename FieldT>
class CondBuilder {
 public:
  CondBuilder(std::vector<std::pair<nebo_detail::NodeP, nebo_detail::NodeP>> c) : clauses_(std::move(c)) {}
  template <typename Test, typename Value>
  CondBuilder<FieldT> operator()(const Test& t, const Value& v) const
  {
    static_assert(nebo_detail::Operand<Test>::boolean, "Nebo: cond test must be a boolean expression or mask");
    auto c = clauses_;
    c.emplace_back(nebo_detail::Operand<Test>::node(t), nebo_detail::Operand<Value>::node(v));
    return CondBuilder<FieldT>(std::move(c));
  }
  template <typename Value>
  nebo_detail::Expr<FieldT> operator()(const Value& v) const
  {
    nebo_detail::NodeP n = nebo_detail::Operand<Value>::node(v);
    for (auto it = clauses_.rbegin(); it != clauses_.rend(); ++it)
      n = nebo_detail::op3(NEBO_COND, it->first, it->second, n);
    return nebo_detail::Expr<FieldT>(n);
  }

 private:
  std::vector<std::pair<nebo_detail::NodeP, nebo_detail::NodeP>> clauses_;
};

template <typename Test, typename Value>
CondBuilder<typename nebo_detail::Common<Test, Value>::type> cond(const Test& t, const Value& v)
{
  static_assert(nebo_detail::Operand<Test>::boolean, "Nebo: cond test must be a boolean expression or mask");
  return CondBuilder<typename nebo_detail::Common<Test, Value>::type>(
      {{nebo_detail::Operand<Test>::node(t), nebo_detail::Operand<Value>::node(v)}});
}

// ---------------------------------------------------------------------------
// assignment (NeboAssignment.h) and reductions (NeboReductions.h)
// ---------------------------------------------------------------------------
namespace nebo_detail {

template <typename FieldT>
void assign(FieldT& lhs, const NodeP& rhs, bool use_ghost)
{
  nebo_field_t d = lhs.descriptor();
  Flat flat(rhs, d.device);
  nebo_expr_t e = flat.expr();
  nebo_check(nebo_assign(&d, &e, use_ghost ? 1 : 0, lhs.get_stream()));
  lhs.mark_written();
}

inline double reduce(const NodeP& expr, int op, bool use_ghost, short device, void* stream)
{
  Flat flat(expr, device);
  nebo_expr_t e = flat.expr();
  double out = 0.0;
  nebo_check(nebo_reduce(&e, op, use_ghost ? 1 : 0, &out, stream));
  return out;
}

template <typename X>
short device_of(const X& x)
{
  Flat probe(Operand<X>::node(x), 0);
  nebo_expr_t e = probe.expr();
  return e.n_fields > 0 ? (short)e.fields[0].device : (short)GPU_INDEX;
}

}  // namespace nebo_detail

// lhs <<= scalar | field | expression (NeboAssignment.h:29-133): returns the
// assigned field as FieldType const&, as the reference does.
template <typename L, typename T, typename R>
typename std::enable_if<nebo_detail::Operand<R>::ok && !nebo_detail::Operand<R>::boolean,
                        const SpatialField<L, T>&>::type
operator<<=(SpatialField<L, T>& lhs, const R& rhs)
{
  typedef typename nebo_detail::Operand<R>::field FR;
  static_assert(std::is_void<FR>::value || std::is_same<FR, SpatialField<L, T>>::value,
                "Nebo: right-hand side field type differs from the left-hand side");
  nebo_detail::assign(lhs, nebo_detail::Operand<R>::node(rhs), true);
  return lhs;
}

template <typename L, typename T, typename R>
const SpatialField<L, T>& interior_assign(SpatialField<L, T>& lhs, const R& rhs)
{
  nebo_detail::assign(lhs, nebo_detail::Operand<R>::node(rhs), false);
  return lhs;
}

namespace nebo_detail {

// lhs(p) = rhs(p) for each interior-relative point of `pts` (x, y, z triples)
template <typename F>
void assign_points(F& lhs, const std::vector<int32_t>& pts, const NodeP& rhs)
{
  nebo_field_t d = lhs.descriptor();
  Flat flat(rhs, d.device);
  nebo_expr_t e = flat.expr();
  nebo_check(nebo_masked_assign(&d, pts.data(), (int32_t)(pts.size() / 3), &e, lhs.get_stream()));
  lhs.mark_written();
}

}  // namespace nebo_detail

template <typename F, typename R>
const F& masked_assign(const SpatialMask<F>& mask, F& lhs, const R& rhs)
{
  nebo_detail::assign_points(lhs, mask.flat_points(), nebo_detail::Operand<R>::node(rhs));
  return lhs;
}

#define NEBO_REDUCTION(NAME, OP, GHOST)                                                         \
  template <typename X>                                                                         \
  double NAME(const X& x)                                                                       \
  {                                                                                             \
    return nebo_detail::reduce(nebo_detail::Operand<X>::node(x), OP, GHOST,                     \
                               nebo_detail::device_of(x), nullptr);                             \
  }
NEBO_REDUCTION(nebo_max, NEBO_RED_MAX, true)
NEBO_REDUCTION(nebo_min, NEBO_RED_MIN, true)
NEBO_REDUCTION(nebo_sum, NEBO_RED_SUM, true)
NEBO_REDUCTION(nebo_norm, NEBO_RED_NORM, true)
NEBO_REDUCTION(nebo_max_interior, NEBO_RED_MAX, false)
NEBO_REDUCTION(nebo_min_interior, NEBO_RED_MIN, false)
NEBO_REDUCTION(nebo_sum_interior, NEBO_RED_SUM, false)
NEBO_REDUCTION(nebo_norm_interior, NEBO_RED_NORM, false)
#undef NEBO_REDUCTION

#define field_max nebo_max
#define field_max_interior nebo_max_interior
#define field_min nebo_min
#define field_min_interior nebo_min_interior
#define field_sum nebo_sum
#define field_sum_interior nebo_sum_interior
#define field_norm nebo_norm
#define field_norm_interior nebo_norm_interior

// periodic copy-fill of the ghost layers (the workloads' boundary handling)
template <typename L, typename T>
void fill_ghosts_periodic(SpatialField<L, T>& f)
{
  nebo_field_t d = f.descriptor();
  nebo_check(nebo_fill_ghosts_periodic(&d, f.get_stream()));
  f.mark_written();
}

}  // namespace SpatialOps

#include <spatialops/NeboStencilBuilder.h>
