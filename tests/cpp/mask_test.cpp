// spatialops/structured/test/testMask.cpp against the C++ front end on the GPU:
// f <<= cond(mask, 3)(mask2, 7)(4) and masked_assign give the literal field.
#include <spatialops/Nebo.h>

#include "nebo_test_helper.h"

using namespace SpatialOps;
typedef SVolField FieldT;

int main()
{
  TestHelper status;
  const GhostData ghost(1);
  const BoundaryCellInfo bcinfo = BoundaryCellInfo::build<FieldT>(false, false, false);
  MemoryWindow w(IntVec(30, 30, 30));
  FieldT f(w, bcinfo, ghost, nullptr, InternalStorage, GPU_INDEX);
  FieldT result(w, bcinfo, ghost, nullptr, InternalStorage, CPU_INDEX);
  for (size_t i = 0; i < w.glob_npts(); ++i) result[i] = 4;
  const int center[4][3] = {{3, 3, 0}, {4, 3, 0}, {3, 4, 0}, {4, 4, 0}};
  const int corner[4][3] = {{1, 1, 0}, {6, 1, 0}, {1, 6, 0}, {6, 6, 0}};
  std::vector<IntVec> maskSet, maskSet2;
  for (auto& p : center) {
    result(p[0] + 1, p[1] + 1, p[2] + 1) = 3;
    maskSet.push_back(IntVec(p[0], p[1], p[2]));
  }
  for (auto& p : corner) {
    result(p[0] + 1, p[1] + 1, p[2] + 1) = 7;
    maskSet2.push_back(IntVec(p[0], p[1], p[2]));
  }
  SpatialMask<FieldT> mask(f, maskSet), mask2(f, maskSet2);
  mask.add_consumer(GPU_INDEX);
  mask2.add_consumer(GPU_INDEX);

  f <<= cond(mask, 3.0)(mask2, 7.0)(4.0);
  f.set_device_as_active(CPU_INDEX);
  bool same = true;
  for (size_t i = 0; i < w.glob_npts(); ++i) same = same && f[i] == result[i];
  status(same, "Cond version");

  f.set_device_as_active(GPU_INDEX);
  f <<= 4.0;
  masked_assign(mask, f, 3.0);
  masked_assign(mask2, f, 7.0);
  f.set_device_as_active(CPU_INDEX);
  same = true;
  for (size_t i = 0; i < w.glob_npts(); ++i) same = same && f[i] == result[i];
  status(same, "masked_assign version");

  // the host-side mask predicate agrees with the point list
  status(mask(3, 3, 0) && mask2(6, 6, 0) && !mask(0, 0, 0), "SpatialMask::operator()");
  return status.report();
}
