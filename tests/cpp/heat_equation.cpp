// examples/heat_equation.cpp of the reference, on the C++ front end and a GPU:
// a 6x6x1 SVolField with masks on the four sides, explicit steps of
//   rhs <<= divX(interpX(alpha)*gradX(phi)) + divY(interpY(alpha)*gradY(phi));
//   phi <<= phi + deltaT*rhs;
//   phi <<= cond(left,10)(right,0)(top||bottom,5)(phi);
// checked against a host loop of the same arithmetic in the same order.
#include <spatialops/Nebo.h>
#include <spatialops/OperatorDatabase.h>

#include <vector>

#include "nebo_test_helper.h"

using namespace SpatialOps;
typedef SVolField FieldT;

int main()
{
  TestHelper status;
  const IntVec dim(6, 6, 1);
  const GhostData nghost(1);
  const BoundaryCellInfo bcInfo = BoundaryCellInfo::build<FieldT>(true, true, true);
  const MemoryWindow window(get_window_with_ghost(dim, nghost, bcInfo));
  FieldT phi(window, bcInfo, nghost, nullptr, InternalStorage, GPU_INDEX);
  FieldT rhs(window, bcInfo, nghost, nullptr, InternalStorage, GPU_INDEX);
  FieldT alpha(window, bcInfo, nghost, nullptr, InternalStorage, GPU_INDEX);
  const Grid grid(dim, {1.0, 1.0, 1.0});
  alpha <<= 1.0;

  std::vector<IntVec> leftSet, rightSet, topSet, bottomSet;
  for (int i = -1; i <= grid.extent(1); ++i) {
    leftSet.push_back(IntVec(-1, i, 0));
    rightSet.push_back(IntVec(grid.extent(0), i, 0));
  }
  for (int i = -1; i <= grid.extent(0); ++i) {
    topSet.push_back(IntVec(i, -1, 0));
    bottomSet.push_back(IntVec(i, grid.extent(1), 0));
  }
  SpatialMask<FieldT> left(phi, leftSet), right(phi, rightSet), top(phi, topSet), bottom(phi, bottomSet);
  left.add_consumer(GPU_INDEX);
  right.add_consumer(GPU_INDEX);
  top.add_consumer(GPU_INDEX);
  bottom.add_consumer(GPU_INDEX);

  OperatorDatabase opDB;
  build_stencils(grid, opDB);
  typedef BasicOpTypes<SVolField> Ops;
  const Ops::DivX& divX = *opDB.retrieve_operator<Ops::DivX>();
  const Ops::GradX& gradX = *opDB.retrieve_operator<Ops::GradX>();
  const Ops::InterpC2FX& interpX = *opDB.retrieve_operator<Ops::InterpC2FX>();
  const Ops::DivY& divY = *opDB.retrieve_operator<Ops::DivY>();
  const Ops::GradY& gradY = *opDB.retrieve_operator<Ops::GradY>();
  const Ops::InterpC2FY& interpY = *opDB.retrieve_operator<Ops::InterpC2FY>();

  const double dx = grid.spacing<XDIR>(), dy = grid.spacing<YDIR>();
  const double deltaT = 0.25 * (dx * dx * dy * dy) / ((dx * dx + dy * dy) * field_min(alpha));

  phi <<= cond(left, 10.0)(right, 0.0)(5.0);

  // host model of the same steps (window 8x8x1, frame origin (1,1,0))
  const int X = window.glob_dim()[0], Y = window.glob_dim()[1];
  std::vector<double> h(X * Y), r(X * Y, 0.0);
  auto at = [&](int x, int y) -> double& { return h[(x + 1) + X * (y + 1)]; };
  auto rat = [&](int x, int y) -> double& { return r[(x + 1) + X * (y + 1)]; };
  for (int y = -1; y <= 6; ++y)
    for (int x = -1; x <= 6; ++x) at(x, y) = x == -1 ? 10.0 : x == 6 ? 0.0 : 5.0;
  const double gx0 = -1.0 / dx, gx1 = 1.0 / dx, gy0 = -1.0 / dy, gy1 = 1.0 / dy;
  for (int step = 0; step <= 40; ++step) {
    rhs <<= divX(interpX(alpha) * gradX(phi)) + divY(interpY(alpha) * gradY(phi));
    phi <<= phi + deltaT * rhs;
    phi <<= cond(left, 10.0)(right, 0.0)(top || bottom, 5.0)(phi);
    // host: rhs over the interior (alpha = 1 everywhere)
    for (int y = 0; y < 6; ++y)
      for (int x = 0; x < 6; ++x) {
        auto fx = [&](int i) { return (1.0 * 0.5 + 1.0 * 0.5) * (at(i - 1, y) * gx0 + at(i, y) * gx1); };
        auto fy = [&](int j) { return (1.0 * 0.5 + 1.0 * 0.5) * (at(x, j - 1) * gy0 + at(x, j) * gy1); };
        rat(x, y) = (fx(x) * gx0 + fx(x + 1) * gx1) + (fy(y) * gy0 + fy(y + 1) * gy1);
      }
    for (int y = -1; y <= 6; ++y)
      for (int x = -1; x <= 6; ++x) at(x, y) = at(x, y) + deltaT * rat(x, y);
    for (int y = -1; y <= 6; ++y)
      for (int x = -1; x <= 6; ++x) {
        if (x == -1) at(x, y) = 10.0;
        else if (x == 6) at(x, y) = 0.0;
        else if (y == -1 || y == 6) at(x, y) = 5.0;
      }
  }
  phi.set_device_as_active(CPU_INDEX);
  bool same = true;
  for (int y = -1; y <= 6; ++y)
    for (int x = -1; x <= 6; ++x) same = same && phi(x + 1, y + 1, 0) == at(x, y);
  status(same, "heat equation: 41 steps equal the host model bit for bit");
  double mx = 0, mn = 1e300;
  for (int y = 0; y < 6; ++y)
    for (int x = 0; x < 6; ++x) {
      mx = std::max(mx, phi(x + 1, y + 1, 0));
      mn = std::min(mn, phi(x + 1, y + 1, 0));
    }
  status(mn >= 0.0 && mx <= 10.0, "heat equation: interior stays within the boundary values");
  return status.report();
}
