// PointFieldTypes.h -- the point field of the C++ front end
// (spatialops/pointfield/PointFieldTypes.h): a SpatialField of (n, 1, 1) points.
#pragma once

#include <spatialops/structured/SpatialField.h>

namespace SpatialOps {
namespace Point {

struct PointFieldTraits {
  typedef NODIR FaceDir;
  static IntVec offset() { return IntVec(0, 0, 0); }
  static IntVec bc_extra() { return IntVec(0, 0, 0); }
  static constexpr int ox = 0, oy = 0, oz = 0;
};

typedef SpatialField<PointFieldTraits> PointField;

}  // namespace Point
}  // namespace SpatialOps
