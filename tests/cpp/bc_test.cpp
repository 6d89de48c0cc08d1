// spatialops/structured/test/testNeboBC.cpp against the C++ front end on the GPU:
// NeboBoundaryConditionBuilder inverts GradX on SVol, GradX on XVol, GradY on
// YVol and the FD GradX on XVol at a minus and a plus mask; the expected values
// are the test's own formulas, compared exactly.
#include <spatialops/structured/FVStaggered.h>

#include "nebo_test_helper.h"

using namespace SpatialOps;

namespace {

const IntVec dim(3, 3, 1);
const double length = 1.5;
const double dx = length / dim[0];
const double dy = length / dim[1];

// a temporary from the store, as the reference test obtains its fields
template <typename F>
SpatFldPtr<F> make()
{
  const BoundaryCellInfo bc = BoundaryCellInfo::build<F>(true, true, true);
  const GhostData g(1);
  return SpatialFieldStore::get_from_window<F>(get_window_with_ghost(dim, g, bc), bc, g);
}

// fill the window through the field iterator with fn(i, j), i and j counted
// from -1 over the window's extent, as the test's loops do
template <typename F, typename Fn>
void fill(F& f, Fn fn)
{
  const MemoryWindow& w = f.window_with_ghost();
  typename F::iterator it = f.begin();
  for (int j = -1; j < (int)w.extent(1) - 1; ++j)
    for (int i = -1; i < (int)w.extent(0) - 1; ++i) *it++ = fn(i, j);
}

template <typename F>
bool equal(F& a, F& b)
{
  a.set_device_as_active(CPU_INDEX);
  const size_t n = a.window_with_ghost().glob_npts();
  bool same = true;
  for (size_t i = 0; i < n; ++i) same = same && a[i] == b[i];
  return same;
}

template <typename Phi, typename Gamma, typename Op, typename Want>
bool run(const std::vector<IntVec>& minusSet, const std::vector<IntVec>& plusSet, Want want)
{
  SpatFldPtr<Phi> test = make<Phi>(), ref = make<Phi>();
  SpatFldPtr<Gamma> gamma = make<Gamma>();
  fill(*test, [](int i, int j) { return double(i + j * 5); });
  fill(*ref, want);
  fill(*gamma, [](int i, int j) { return double(25 + i + j * 5); });
  test->set_device_as_active(GPU_INDEX);
  gamma->set_device_as_active(GPU_INDEX);

  OperatorDatabase opdb;
  build_stencils(dim[0], dim[1], dim[2], length, length, length, opdb);
  const Op* const op = opdb.retrieve_operator<Op>();
  NeboBoundaryConditionBuilder<Op> BC(*op);

  SpatialMask<Gamma> minus(*gamma, minusSet);
  minus.add_consumer(GPU_INDEX);
  BC(minus, *test, *gamma, true);
  SpatialMask<Gamma> plus = SpatialMask<Gamma>::build(*test, plusSet);
  plus.add_consumer(GPU_INDEX);
  BC(plus, *test, *gamma, false);
  return equal(*test, *ref);
}

}  // namespace

int main()
{
  TestHelper status;

  status(run<SVolField, SSurfXField, BasicOpTypes<SVolField>::GradX>(
             {IntVec(0, 1, 0), IntVec(1, 2, 0)}, {IntVec(2, 1, 0), IntVec(3, 2, 0)},
             [](int i, int j) {
               if ((i == -1 && j == 1) || (i == 0 && j == 2)) {
                 int B = 25 + (i + 1) + j * 5, C = (i + 1) + j * 5;
                 return (B - C * (1 / dx)) / (-1 / dx);
               }
               if ((i == 2 && j == 1) || (i == 3 && j == 2)) {
                 int B = 25 + i + j * 5, A = (i - 1) + j * 5;
                 return (B - A * (-1 / dx)) / (1 / dx);
               }
               return double(i + j * 5);
             }),
         "Invert GradX on SVol");

  status(run<XVolField, XSurfXField, BasicOpTypes<XVolField>::GradX>(
             {IntVec(-1, 1, 0), IntVec(-1, 2, 0)}, {IntVec(3, 1, 0), IntVec(3, 2, 0)},
             [](int i, int j) {
               if ((i == -1 && j == 1) || (i == -1 && j == 2)) {
                 int B = 25 + i + j * 5, C = (i + 1) + j * 5;
                 return (B - C * (1 / dx)) / (-1 / dx);
               }
               if ((i == 4 && j == 1) || (i == 4 && j == 2)) {
                 int B = 25 + (i - 1) + j * 5, A = (i - 1) + j * 5;
                 return (B - A * (-1 / dx)) / (1 / dx);
               }
               return double(i + j * 5);
             }),
         "Invert GradX on XVol");

  status(run<YVolField, YSurfYField, BasicOpTypes<YVolField>::GradY>(
             {IntVec(0, -1, 0), IntVec(1, -1, 0)}, {IntVec(0, 3, 0), IntVec(3, 3, 0)},
             [](int i, int j) {
               if ((i == 0 && j == -1) || (i == 1 && j == -1)) {
                 int B = 25 + i + j * 5, C = i + (j + 1) * 5;
                 return (B - C * (1 / dy)) / (-1 / dy);
               }
               if ((i == 0 && j == 4) || (i == 3 && j == 4)) {
                 int B = 25 + i + (j - 1) * 5, A = i + (j - 1) * 5;
                 return (B - A * (-1 / dy)) / (1 / dy);
               }
               return double(i + j * 5);
             }),
         "Invert GradY on YVol");

  status(run<XVolField, XVolField, OperatorTypeBuilder<GradientX, XVolField, XVolField>::type>(
             {IntVec(0, 0, 0), IntVec(0, 1, 0)}, {IntVec(3, 0, 0), IntVec(3, 1, 0)},
             [](int i, int j) {
               if ((i == -1 && j == 0) || (i == -1 && j == 1)) {
                 int B = 25 + (i + 1) + j * 5, C = (i + 2) + j * 5;
                 return (B - C * (0.5 / dx)) / (-0.5 / dx);
               }
               if ((i == 4 && j == 0) || (i == 4 && j == 1)) {
                 int B = 25 + (i - 1) + j * 5, A = (i - 2) + j * 5;
                 return (B - A * (-0.5 / dx)) / (0.5 / dx);
               }
               return double(i + j * 5);
             }),
         "Invert GradX(FD) on XVol");

  return status.report();
}
