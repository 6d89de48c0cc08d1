// NeboStencilBuilder.h -- stencil operators of the C++ front end: the builders of
// spatialops/NeboStencilBuilder.h, the operator types of
// spatialops/structured/stencil/FVStaggeredOperatorTypes.h, OperatorDatabase
// (spatialops/OperatorDatabase.h) and build_stencils
// (spatialops/structured/stencil/StencilBuilder.cpp), plus Grid.
//
// Point sets come from the locations' staggering offsets: Stencil2 reads
// -(src > dest) and (src < dest) (NeboStencilBuilder.h:154-166), Stencil4 pairs
// a low/high value in two directions (:168-221), FD stencils read -e_d and +e_d
// (:223-232).  Coefficients follow StencilBuilder.cpp:60-72.
#pragma once

#include <spatialops/Nebo.h>

#include <typeindex>

namespace SpatialOps {

// operator type tags (SpatialOpsDefs.h)
struct Interpolant {};
struct Gradient {};
struct Divergence {};
struct Filter {};
struct InterpolantX { typedef XDIR DirT; };
struct InterpolantY { typedef YDIR DirT; };
struct InterpolantZ { typedef ZDIR DirT; };
struct GradientX { typedef XDIR DirT; };
struct GradientY { typedef YDIR DirT; };
struct GradientZ { typedef ZDIR DirT; };

namespace nebo_detail {

typedef std::vector<std::array<int, 3>> Points;

template <typename Src, typename Dest>
Points stencil2_points()
{
  const IntVec s = Src::Location::offset(), d = Dest::Location::offset();
  std::array<int, 3> lo, hi;
  for (int a = 0; a < 3; ++a) {
    lo[a] = -(s[a] > d[a] ? 1 : 0);
    hi[a] = s[a] < d[a] ? 1 : 0;
  }
  return {lo, hi};
}

template <typename Src, typename Dest>
Points stencil4_points()
{
  const IntVec s = Src::Location::offset(), d = Dest::Location::offset();
  const int first = s[0] != d[0] ? 0 : 1;
  const int second = s[2] != d[2] ? 2 : 1;
  auto lo = [&](int a) { return -(s[a] > d[a] ? 1 : 0); };
  auto hi = [&](int a) { return s[a] < d[a] ? 1 : 0; };
  auto pt = [&](int f, int g) {
    std::array<int, 3> p = {0, 0, 0};
    p[first] += f;
    p[second] += g;
    return p;
  };
  return {pt(lo(first), lo(second)), pt(hi(first), lo(second)), pt(lo(first), hi(second)),
          pt(hi(first), hi(second))};
}

inline Points fd_points(int axis)
{
  std::array<int, 3> lo = {0, 0, 0}, hi = {0, 0, 0};
  lo[axis] = -1;
  hi[axis] = 1;
  return {lo, hi};
}

inline Points box27_points()
{
  Points p;
  for (int z = -1; z <= 1; ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) p.push_back({x, y, z});
  return p;
}

}  // namespace nebo_detail

// Weighted stencil: sum_p arg(x+p)*c_p, folded left in point order.
template <typename OpT, typename SrcT, typename DestT>
class NeboStencilBuilder {
 public:
  typedef OpT type;
  typedef SrcT SrcFieldType;
  typedef DestT DestFieldType;

  NeboStencilBuilder(nebo_detail::Points pts, std::vector<double> coefs)
  {
    auto s = std::make_shared<nebo_detail::StencilSpec>();
    s->kind = NEBO_STENCIL_WEIGHTED;
    s->pts = std::move(pts);
    s->coefs = std::move(coefs);
    if (s->coefs.size() != s->pts.size())
      throw std::runtime_error("Nebo error in Nebo Stencil Coefficient Collection:\n"
                               "trying to access a coefficient that does not exist");
    spec_ = s;
  }
  nebo_detail::Expr<DestT> operator()(const SrcT& src) const
  {
    return nebo_detail::Expr<DestT>(nebo_detail::stencil(spec_, nebo_detail::field_leaf(src)));
  }
  nebo_detail::Expr<DestT> operator()(const nebo_detail::Expr<SrcT>& src) const
  {
    return nebo_detail::Expr<DestT>(nebo_detail::stencil(spec_, src.node()));
  }
  void apply_to_field(const SrcT& src, DestT& dest) const { dest <<= (*this)(src); }
  const std::vector<double>& coefs() const { return spec_->coefs; }
  const nebo_detail::Points& points() const { return spec_->pts; }

 private:
  std::shared_ptr<const nebo_detail::StencilSpec> spec_;
};

// Plain sum stencil (NeboSumStencil) and its average (NeboAverageStencilBuilder).
template <typename SrcT, typename DestT>
class NeboSumStencilBuilder {
 public:
  typedef SrcT SrcFieldType;
  typedef DestT DestFieldType;
  explicit NeboSumStencilBuilder(nebo_detail::Points pts, double scale = 0.0) : scale_(scale)
  {
    auto s = std::make_shared<nebo_detail::StencilSpec>();
    s->kind = NEBO_STENCIL_SUM;
    s->pts = std::move(pts);
    spec_ = s;
  }
  nebo_detail::Expr<DestT> operator()(const SrcT& src) const { return wrap(nebo_detail::field_leaf(src)); }
  nebo_detail::Expr<DestT> operator()(const nebo_detail::Expr<SrcT>& src) const { return wrap(src.node()); }
  void apply_to_field(const SrcT& src, DestT& dest) const { dest <<= (*this)(src); }

 private:
  nebo_detail::Expr<DestT> wrap(nebo_detail::NodeP arg) const
  {
    nebo_detail::NodeP n = nebo_detail::stencil(spec_, std::move(arg));
    if (scale_ != 0.0) n = nebo_detail::op2(NEBO_MUL, n, nebo_detail::scalar(scale_));
    return nebo_detail::Expr<DestT>(n);
  }
  std::shared_ptr<const nebo_detail::StencilSpec> spec_;
  double scale_;
};

// Box filter: the 27-point sum times 1/27 (BoxFilter3DStencilCollection).
template <typename SrcT, typename DestT>
class NeboAverageStencilBuilder : public NeboSumStencilBuilder<SrcT, DestT> {
 public:
  NeboAverageStencilBuilder() : NeboSumStencilBuilder<SrcT, DestT>(nebo_detail::box27_points(), 1.0 / 27.0) {}
};

// OperatorTypeBuilder<Op, Src, Dest>::type : the operator class for a triple.
template <typename OpT, typename SrcT, typename DestT>
struct OperatorTypeBuilder { typedef NeboStencilBuilder<OpT, SrcT, DestT> type; };
template <typename FieldT>
struct OperatorTypeBuilder<Interpolant, FieldT, FieldT> { typedef NeboSumStencilBuilder<FieldT, FieldT> type; };
template <typename FieldT>
struct OperatorTypeBuilder<Filter, FieldT, FieldT> { typedef NeboAverageStencilBuilder<FieldT, FieldT> type; };

template <typename VolT>
struct BasicOpTypes {
  typedef typename FaceTypes<VolT>::XFace XF;
  typedef typename FaceTypes<VolT>::YFace YF;
  typedef typename FaceTypes<VolT>::ZFace ZF;
  typedef typename OperatorTypeBuilder<Interpolant, VolT, XF>::type InterpC2FX;
  typedef typename OperatorTypeBuilder<Interpolant, VolT, YF>::type InterpC2FY;
  typedef typename OperatorTypeBuilder<Interpolant, VolT, ZF>::type InterpC2FZ;
  typedef typename OperatorTypeBuilder<Interpolant, XF, VolT>::type InterpF2CX;
  typedef typename OperatorTypeBuilder<Interpolant, YF, VolT>::type InterpF2CY;
  typedef typename OperatorTypeBuilder<Interpolant, ZF, VolT>::type InterpF2CZ;
  typedef typename OperatorTypeBuilder<Gradient, VolT, XF>::type GradX;
  typedef typename OperatorTypeBuilder<Gradient, VolT, YF>::type GradY;
  typedef typename OperatorTypeBuilder<Gradient, VolT, ZF>::type GradZ;
  typedef typename OperatorTypeBuilder<Divergence, XF, VolT>::type DivX;
  typedef typename OperatorTypeBuilder<Divergence, YF, VolT>::type DivY;
  typedef typename OperatorTypeBuilder<Divergence, ZF, VolT>::type DivZ;
};

// ---------------------------------------------------------------------------
// OperatorDatabase (OperatorDatabase.h:56-190): one operator per type
// ---------------------------------------------------------------------------
class OperatorDatabase {
 public:
  template <typename OpT>
  OpT* register_new_operator(OpT* op)
  {
    ops_[std::type_index(typeid(OpT))] = std::shared_ptr<void>(op, [](void* p) { delete static_cast<OpT*>(p); });
    return op;
  }
  template <typename OpT>
  OpT* retrieve_operator() const
  {
    auto it = ops_.find(std::type_index(typeid(OpT)));
    if (it == ops_.end())
      throw std::runtime_error(std::string("Nebo error in OperatorDatabase:\nno operator of type ") +
                               typeid(OpT).name() + " registered");
    return static_cast<OpT*>(it->second.get());
  }
  template <typename OpT>
  bool has_operator() const { return ops_.count(std::type_index(typeid(OpT))) != 0; }
  size_t size() const { return ops_.size(); }

 private:
  std::map<std::type_index, std::shared_ptr<void>> ops_;
};

namespace nebo_detail {

template <typename Op, typename Src, typename Dest>
void reg2(OperatorDatabase& db, std::vector<double> c)
{
  db.register_new_operator(new NeboStencilBuilder<Op, Src, Dest>(stencil2_points<Src, Dest>(), std::move(c)));
}

template <typename Src, typename Dest>
void reg4(OperatorDatabase& db)
{
  db.register_new_operator(
      new NeboStencilBuilder<Interpolant, Src, Dest>(stencil4_points<Src, Dest>(), {0.25, 0.25, 0.25, 0.25}));
}

template <typename Src, typename Dest>
void reg_null(OperatorDatabase& db)
{
  db.register_new_operator(new NeboSumStencilBuilder<Src, Dest>({{0, 0, 0}}));
}

template <typename F>
void reg_filter(OperatorDatabase& db)
{
  db.register_new_operator(new typename OperatorTypeBuilder<Filter, F, F>::type());
}

template <typename Vol>
void reg_basic(OperatorDatabase& db, const double dcoef[3][2])
{
  typedef FaceTypes<Vol> FT;
  const std::vector<double> half = {0.5, 0.5};
  reg2<Interpolant, Vol, typename FT::XFace>(db, half);
  reg2<Interpolant, Vol, typename FT::YFace>(db, half);
  reg2<Interpolant, Vol, typename FT::ZFace>(db, half);
  reg2<Interpolant, typename FT::XFace, Vol>(db, half);
  reg2<Interpolant, typename FT::YFace, Vol>(db, half);
  reg2<Interpolant, typename FT::ZFace, Vol>(db, half);
  reg2<Gradient, Vol, typename FT::XFace>(db, {dcoef[0][0], dcoef[0][1]});
  reg2<Gradient, Vol, typename FT::YFace>(db, {dcoef[1][0], dcoef[1][1]});
  reg2<Gradient, Vol, typename FT::ZFace>(db, {dcoef[2][0], dcoef[2][1]});
  reg2<Divergence, typename FT::XFace, Vol>(db, {dcoef[0][0], dcoef[0][1]});
  reg2<Divergence, typename FT::YFace, Vol>(db, {dcoef[1][0], dcoef[1][1]});
  reg2<Divergence, typename FT::ZFace, Vol>(db, {dcoef[2][0], dcoef[2][1]});
}

template <typename Op, int Axis, typename Vol>
void reg_fd(OperatorDatabase& db, std::vector<double> c)
{
  db.register_new_operator(new NeboStencilBuilder<Op, Vol, Vol>(fd_points(Axis), std::move(c)));
}

template <typename Vol>
void reg_fd_all(OperatorDatabase& db, const double hc[3][2])
{
  const std::vector<double> half = {0.5, 0.5};
  reg_fd<InterpolantX, 0, Vol>(db, half);
  reg_fd<InterpolantY, 1, Vol>(db, half);
  reg_fd<InterpolantZ, 2, Vol>(db, half);
  reg_fd<GradientX, 0, Vol>(db, {hc[0][0], hc[0][1]});
  reg_fd<GradientY, 1, Vol>(db, {hc[1][0], hc[1][1]});
  reg_fd<GradientZ, 2, Vol>(db, {hc[2][0], hc[2][1]});
}

}  // namespace nebo_detail

// build_stencils (StencilBuilder.cpp:52-220)
inline void build_stencils(unsigned nx, unsigned ny, unsigned nz, double Lx, double Ly, double Lz,
                           OperatorDatabase& db)
{
  using namespace nebo_detail;
  const double dx = Lx / nx, dy = Ly / ny, dz = Lz / nz;
  const double dc[3][2] = {{-1.0 / dx, 1.0 / dx}, {-1.0 / dy, 1.0 / dy}, {-1.0 / dz, 1.0 / dz}};
  const double hc[3][2] = {{-0.5 / dx, 0.5 / dx}, {-0.5 / dy, 0.5 / dy}, {-0.5 / dz, 0.5 / dz}};
  const std::vector<double> half = {0.5, 0.5};
  reg_basic<SVolField>(db, dc);
  reg_basic<XVolField>(db, dc);
  reg_basic<YVolField>(db, dc);
  reg_basic<ZVolField>(db, dc);
  reg2<Interpolant, XVolField, YSurfXField>(db, half);
  reg2<Gradient, XVolField, YSurfXField>(db, {dc[1][0], dc[1][1]});
  reg2<Interpolant, XVolField, ZSurfXField>(db, half);
  reg2<Gradient, XVolField, ZSurfXField>(db, {dc[2][0], dc[2][1]});
  reg2<Interpolant, YVolField, XSurfYField>(db, half);
  reg2<Gradient, YVolField, XSurfYField>(db, {dc[0][0], dc[0][1]});
  reg2<Interpolant, YVolField, ZSurfYField>(db, half);
  reg2<Gradient, YVolField, ZSurfYField>(db, {dc[2][0], dc[2][1]});
  reg2<Interpolant, ZVolField, XSurfZField>(db, half);
  reg2<Gradient, ZVolField, XSurfZField>(db, {dc[0][0], dc[0][1]});
  reg2<Interpolant, ZVolField, YSurfZField>(db, half);
  reg2<Gradient, ZVolField, YSurfZField>(db, {dc[1][0], dc[1][1]});
  reg2<Interpolant, SVolField, XVolField>(db, half);
  reg2<Gradient, SVolField, XVolField>(db, {dc[0][0], dc[0][1]});
  reg2<Interpolant, SVolField, YVolField>(db, half);
  reg2<Gradient, SVolField, YVolField>(db, {dc[1][0], dc[1][1]});
  reg2<Interpolant, SVolField, ZVolField>(db, half);
  reg2<Gradient, SVolField, ZVolField>(db, {dc[2][0], dc[2][1]});
  reg2<Interpolant, XVolField, SVolField>(db, half);
  reg2<Gradient, XVolField, SVolField>(db, {dc[0][0], dc[0][1]});
  reg2<Interpolant, YVolField, SVolField>(db, half);
  reg2<Gradient, YVolField, SVolField>(db, {dc[1][0], dc[1][1]});
  reg2<Interpolant, ZVolField, SVolField>(db, half);
  reg2<Gradient, ZVolField, SVolField>(db, {dc[2][0], dc[2][1]});
  reg_null<SVolField, SVolField>(db);
  reg_null<XVolField, XVolField>(db);
  reg_null<YVolField, YVolField>(db);
  reg_null<ZVolField, ZVolField>(db);
  reg_null<XVolField, SSurfXField>(db);
  reg_null<YVolField, SSurfYField>(db);
  reg_null<ZVolField, SSurfZField>(db);
  reg_null<SVolField, XSurfXField>(db);
  reg_null<SVolField, YSurfYField>(db);
  reg_null<SVolField, ZSurfZField>(db);
  reg_null<XSurfXField, SVolField>(db);
  reg_null<YSurfYField, SVolField>(db);
  reg_null<ZSurfZField, SVolField>(db);
  reg4<SVolField, XSurfYField>(db);
  reg4<SVolField, XSurfZField>(db);
  reg4<SVolField, YSurfXField>(db);
  reg4<SVolField, YSurfZField>(db);
  reg4<SVolField, ZSurfXField>(db);
  reg4<SVolField, ZSurfYField>(db);
  reg4<XSurfYField, SVolField>(db);
  reg4<XSurfZField, SVolField>(db);
  reg4<YSurfXField, SVolField>(db);
  reg4<YSurfZField, SVolField>(db);
  reg4<ZSurfXField, SVolField>(db);
  reg4<ZSurfYField, SVolField>(db);
  reg4<XVolField, YVolField>(db);
  reg4<XVolField, ZVolField>(db);
  reg4<YVolField, XVolField>(db);
  reg4<YVolField, ZVolField>(db);
  reg4<ZVolField, XVolField>(db);
  reg4<ZVolField, YVolField>(db);
  reg_filter<SVolField>(db);
  reg_filter<XVolField>(db);
  reg_filter<YVolField>(db);
  reg_filter<ZVolField>(db);
  reg_filter<XSurfXField>(db);
  reg_filter<XSurfYField>(db);
  reg_filter<XSurfZField>(db);
  reg_filter<YSurfXField>(db);
  reg_filter<YSurfYField>(db);
  reg_filter<YSurfZField>(db);
  reg_filter<ZSurfXField>(db);
  reg_filter<ZSurfYField>(db);
  reg_filter<ZSurfZField>(db);
  reg_fd_all<SVolField>(db, hc);
  reg_fd_all<XVolField>(db, hc);
  reg_fd_all<YVolField>(db, hc);
  reg_fd_all<ZVolField>(db, hc);
}

// ---------------------------------------------------------------------------
// Boundary conditions (NeboStencilBuilder.h:504-764)
// ---------------------------------------------------------------------------
namespace nebo_detail {

// Direction of an operator type tag (GradientX -> 0, ...), -1 when it has none.
template <typename Op, typename = void>
struct op_axis { static constexpr int value = -1; };
template <typename Op>
struct op_axis<Op, typename std::conditional<true, void, typename Op::DirT>::type> {
  static constexpr int value = Op::DirT::value;
};

template <typename OpT, typename Src, typename Dest>
struct MaskShiftPoints {
  // MaskShiftPoints (:504-516): minus = (src < dest), plus = -(src > dest) per
  // component of the location offsets; a same-location (FD) operator shifts by
  // +e and -e along its direction (:518-526).
  static std::array<int, 3> get(bool minus)
  {
    std::array<int, 3> p = {0, 0, 0};
    if (std::is_same<Src, Dest>::value) {
      static_assert(!std::is_same<Src, Dest>::value || op_axis<OpT>::value >= 0,
                    "Nebo: a mask shift between fields of one type needs a directional operator");
      p[op_axis<OpT>::value >= 0 ? op_axis<OpT>::value : 0] = minus ? 1 : -1;
      return p;
    }
    const IntVec s = Src::Location::offset(), d = Dest::Location::offset();
    for (int a = 0; a < 3; ++a) p[a] = minus ? (s[a] < d[a] ? 1 : 0) : -(s[a] > d[a] ? 1 : 0);
    return p;
  }
};

}  // namespace nebo_detail

// NeboMaskShiftBuilder<Op, Src, Dest>: a mask of Src points as a boolean
// expression over Dest points, mask(x + shift).
template <typename OperatorT, typename SrcFieldT, typename DestFieldT>
struct NeboMaskShiftBuilder {
  typedef OperatorT OperatorType;
  typedef SrcFieldT SrcFieldType;
  typedef DestFieldT DestFieldType;

  static std::array<int, 3> minus_point()
  {
    return nebo_detail::MaskShiftPoints<typename OperatorT::type, SrcFieldT, DestFieldT>::get(true);
  }
  static std::array<int, 3> plus_point()
  {
    return nebo_detail::MaskShiftPoints<typename OperatorT::type, SrcFieldT, DestFieldT>::get(false);
  }
  nebo_detail::BoolExpr<DestFieldT> minus(const SpatialMask<SrcFieldT>& src) const
  {
    return nebo_detail::BoolExpr<DestFieldT>(nebo_detail::mask_shift(src.node(), minus_point()));
  }
  nebo_detail::BoolExpr<DestFieldT> plus(const SpatialMask<SrcFieldT>& src) const
  {
    return nebo_detail::BoolExpr<DestFieldT>(nebo_detail::mask_shift(src.node(), plus_point()));
  }
};

// NeboBoundaryConditionBuilder<Op>(op)(mask, phi, gamma, minus): sets, for each
// gamma point g of the mask, the phi value at x = g - shift so that op(phi)(g)
// equals gamma(g).  The low point is op's first point (coefficient coefs()[0]),
// the high point its last:
//   phi(x) = (gamma(x - p_side) * 1.0 - fold_{i != side} phi(x + p_i - p_side) * c_i) / c_side
// with edgeless stencils.  The reference's CPU path loops once over the mask
// points (:646-680); this runs that loop as one masked assignment on the GPU,
// touching only the boundary points (its GPU path sweeps the whole field with a
// cond, :712-719, which gives the same values).
template <typename OperatorT>
struct NeboBoundaryConditionBuilder {
  typedef OperatorT OperatorType;
  typedef typename OperatorT::SrcFieldType PhiFieldType;
  typedef typename OperatorT::DestFieldType GammaFieldType;
  typedef NeboMaskShiftBuilder<OperatorT, GammaFieldType, PhiFieldType> Shift;

  explicit NeboBoundaryConditionBuilder(const OperatorT& op)
  {
    const nebo_detail::Points& p = op.points();
    const std::vector<double>& c = op.coefs();
    if (p.size() < 2)
      throw std::runtime_error("Nebo error in Nebo Boundary Condition:\n"
                               "the operator needs two or more points");
    lowCoef_ = c.front();
    highCoef_ = c.back();
    auto sub = [](const std::array<int, 3>& a, const std::array<int, 3>& b) {
      return std::array<int, 3>{a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    };
    const std::array<int, 3> zero = {0, 0, 0};
    minusGamma_ = edgeless({sub(zero, p.front())}, {1.0});
    plusGamma_ = edgeless({sub(zero, p.back())}, {1.0});
    nebo_detail::Points mp, pp;
    for (size_t i = 1; i < p.size(); ++i) mp.push_back(sub(p[i], p.front()));
    for (size_t i = 0; i + 1 < p.size(); ++i) pp.push_back(sub(p[i], p.back()));
    minusPhi_ = edgeless(mp, std::vector<double>(c.begin() + 1, c.end()));
    plusPhi_ = edgeless(pp, std::vector<double>(c.begin(), c.end() - 1));
  }

  void operator()(const SpatialMask<GammaFieldType>& mask, PhiFieldType& phi,
                  const nebo_detail::Expr<GammaFieldType>& gamma, bool minus) const
  {
    apply(mask, phi, gamma.node(), minus);
  }
  void operator()(const SpatialMask<GammaFieldType>& mask, PhiFieldType& phi,
                  const GammaFieldType& gamma, bool minus) const
  {
    apply(mask, phi, nebo_detail::field_leaf(gamma), minus);
  }
  void operator()(const SpatialMask<GammaFieldType>& mask, PhiFieldType& phi, double gamma,
                  bool minus) const
  {
    apply(mask, phi, nebo_detail::scalar(gamma), minus);
  }

  // The inverted value at phi points, (gamma term - phi term) / coefficient.
  nebo_detail::Expr<PhiFieldType> rhs(const PhiFieldType& phi, const nebo_detail::NodeP& gamma,
                                   bool minus) const
  {
    using namespace nebo_detail;
    NodeP g = stencil(minus ? minusGamma_ : plusGamma_, gamma);
    NodeP f = stencil(minus ? minusPhi_ : plusPhi_, field_leaf(phi));
    return nebo_detail::Expr<PhiFieldType>(
        op2(NEBO_DIV, op2(NEBO_SUB, g, f), scalar(minus ? lowCoef_ : highCoef_)));
  }

 private:
  static std::shared_ptr<const nebo_detail::StencilSpec> edgeless(nebo_detail::Points p,
                                                                  std::vector<double> c)
  {
    auto s = std::make_shared<nebo_detail::StencilSpec>();
    s->kind = NEBO_STENCIL_EDGELESS;
    s->pts = std::move(p);
    s->coefs = std::move(c);
    return s;
  }
  void apply(const SpatialMask<GammaFieldType>& mask, PhiFieldType& phi, const nebo_detail::NodeP& gamma,
             bool minus) const
  {
    const std::array<int, 3> sh = minus ? Shift::minus_point() : Shift::plus_point();
    std::vector<int32_t> pts(mask.flat_points());
    for (size_t i = 0; i < pts.size(); ++i) pts[i] -= sh[i % 3];
    nebo_detail::assign_points(phi, pts, rhs(phi, gamma, minus).node());
  }

  double lowCoef_ = 0.0, highCoef_ = 0.0;
  std::shared_ptr<const nebo_detail::StencilSpec> minusGamma_, plusGamma_, minusPhi_, plusPhi_;
};

// Grid (spatialops/structured/Grid.h): extents, lengths, spacings, coordinates
class Grid {
 public:
  Grid(const IntVec& npts, const std::array<double, 3>& length) : n_(npts), L_(length) {}
  int extent(int i) const { return n_[i]; }
  const IntVec& extent() const { return n_; }
  double length(int i) const { return L_[i]; }
  template <typename Dir>
  double spacing() const { return L_[Dir::value] / n_[Dir::value]; }
  // cell-centre coordinate along Dir over the whole window (ghosts included):
  // dx*(i - ghost) + dx/2*(1 + stagger) with stagger the location offset
  template <typename Dir, typename FieldT>
  void set_coord(FieldT& f) const
  {
    const int d = Dir::value;
    const double dx = L_[d] / n_[d];
    const MemoryWindow& w = f.window_with_ghost();
    const int off = FieldT::Location::offset()[d];
    double* v = f.field_values(CPU_INDEX);
    const int g = f.get_ghost_data().get_minus(d);
    for (int k = 0; k < (int)w.extent(2); ++k)
      for (int j = 0; j < (int)w.extent(1); ++j)
        for (int i = 0; i < (int)w.extent(0); ++i) {
          const int idx[3] = {i, j, k};
          v[w.flat_index(IntVec(i, j, k))] = dx * (idx[d] - g) + 0.5 * dx * (1 + off);
        }
  }

 private:
  IntVec n_;
  std::array<double, 3> L_;
};

inline void build_stencils(const Grid& g, OperatorDatabase& db)
{
  build_stencils(g.extent(0), g.extent(1), g.extent(2), g.length(0), g.length(1), g.length(2), db);
}

}  // namespace SpatialOps
