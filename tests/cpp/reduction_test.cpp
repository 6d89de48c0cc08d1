// field_max / field_min / field_sum / field_norm (and _interior) of the C++ front
// end against host folds (spatialops/NeboReductions.h:45-240).
#include <spatialops/Nebo.h>

#include <algorithm>
#include <cmath>

#include "nebo_test_helper.h"

using namespace SpatialOps;
typedef SVolField FieldT;

int main()
{
  TestHelper status;
  const GhostData ghost(1);
  const BoundaryCellInfo bc = BoundaryCellInfo::build<FieldT>(false, false, false);
  const MemoryWindow w(get_window_with_ghost(IntVec(33, 20, 9), ghost, bc));
  FieldT f(w, bc, ghost, nullptr, InternalStorage, CPU_INDEX);
  unsigned s = 12345u;
  for (size_t i = 0; i < w.glob_npts(); ++i) {
    s = s * 1664525u + 1013904223u;
    f[i] = (double)(s >> 8) / (1u << 24) - 0.5;
  }
  double mx = -1e300, mn = 1e300, sum = 0.0, sq = 0.0, imx = -1e300;
  for (size_t i = 0; i < w.glob_npts(); ++i) {
    mx = std::max(mx, f[i]);
    mn = std::min(mn, f[i]);
    sum += f[i];
    sq += f[i] * f[i];
  }
  for (int k = 1; k <= 9; ++k)
    for (int j = 1; j <= 20; ++j)
      for (int i = 1; i <= 33; ++i) imx = std::max(imx, f(i, j, k));
  f.set_device_as_active(GPU_INDEX);
  status(field_max(f) == mx, "field_max");
  status(field_min(f) == mn, "field_min");
  status(std::fabs(field_sum(f) - sum) <= 1e-12, "field_sum");
  status(std::fabs(field_norm(f) - std::sqrt(sq)) <= 1e-12, "field_norm");
  status(field_max_interior(f) == imx, "field_max_interior");
  status(field_max(f * 2.0 + 1.0) == 2.0 * mx + 1.0, "field_max of an expression");
  return status.report();
}
