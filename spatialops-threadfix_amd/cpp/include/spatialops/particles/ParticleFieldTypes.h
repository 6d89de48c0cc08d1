// ParticleFieldTypes.h -- the particle field of the C++ front end
// (spatialops/particles/ParticleFieldTypes.h): a SpatialField whose window is
// (number of particles, 1, 1), no ghosts, no staggering.
#pragma once

#include <spatialops/structured/SpatialField.h>

namespace SpatialOps {
namespace Particle {

struct ParticleFieldTraits {
  typedef NODIR FaceDir;
  static IntVec offset() { return IntVec(0, 0, 0); }
  static IntVec bc_extra() { return IntVec(0, 0, 0); }
  static constexpr int ox = 0, oy = 0, oz = 0;
};

typedef SpatialField<ParticleFieldTraits> ParticleField;

}  // namespace Particle
}  // namespace SpatialOps
