// ParticleOperators.h -- ParticleToCell and CellToParticle of the C++ front end
// (spatialops/particles/ParticleOperators.h:35-395) on the engine's GPU kernels
// (nebo_particle_to_cell / nebo_cell_to_particle): the same construction,
// set_coordinate_information and apply_to_field, the same values bit for bit.
// The cell field and every particle field must have valid copies on one GPU.
#pragma once

#include <spatialops/particles/ParticleFieldTypes.h>
#include <spatialops/Nebo.h>

namespace SpatialOps {
namespace Particle {

namespace detail {

class ParticleOpBase {
 public:
  ParticleOpBase(double dx, double xlo, double dy, double ylo, double dz, double zlo)
  {
    geom_.dx = dx;
    geom_.xlo = xlo;
    geom_.dy = dy;
    geom_.ylo = ylo;
    geom_.dz = dz;
    geom_.zlo = zlo;
  }
  void set_coordinate_information(const ParticleField* px, const ParticleField* py, const ParticleField* pz,
                                  const ParticleField* psize)
  {
    px_ = px;
    py_ = py;
    pz_ = pz;
    psize_ = psize;
  }

 protected:
  static const double* on(const ParticleField* f, short dev)
  {
    return f ? f->field_values(dev) : nullptr;
  }
  void check(size_t np) const
  {
    if (!psize_)
      throw std::runtime_error("Nebo error in Nebo Particle Interpolation:\nset_coordinate_information first");
    for (const ParticleField* f : {px_, py_, pz_, psize_})
      if (f && f->window_with_ghost().glob_npts() < np)
        throw std::runtime_error(
            "Nebo error in Nebo Particle Interpolation:\na particle field is shorter than the particle count");
  }
  nebo_particle_geom_t geom_;
  const ParticleField *px_ = nullptr, *py_ = nullptr, *pz_ = nullptr, *psize_ = nullptr;
};

}  // namespace detail

// Spreads an extensive particle quantity over the cells each particle overlaps.
template <typename CellField>
class ParticleToCell : public detail::ParticleOpBase {
 public:
  typedef ParticleField SrcFieldType;
  typedef CellField DestFieldType;
  ParticleToCell(double dx, double xlo, double dy = -1, double ylo = 0, double dz = -1, double zlo = 0)
      : ParticleOpBase(dx, xlo, dy, ylo, dz, zlo) {}

  void apply_to_field(const SrcFieldType& src, DestFieldType& dest) const
  {
    const size_t np = src.window_with_ghost().glob_npts();
    check(np);
    nebo_field_t d = dest.descriptor();
    const short dev = (short)d.device;
    nebo_check(nebo_particle_to_cell(&d, src.field_values(dev), on(px_, dev), on(py_, dev), on(pz_, dev),
                                     on(psize_, dev), (int64_t)np, &geom_, nullptr, dest.get_stream()));
    dest.mark_written();
  }
};

// The volume-weighted average of a cell field over each particle's cube.
template <typename CellField>
class CellToParticle : public detail::ParticleOpBase {
 public:
  typedef CellField SrcFieldType;
  typedef ParticleField DestFieldType;
  CellToParticle(double dx, double xlo, double dy = -1, double ylo = 0, double dz = -1, double zlo = 0)
      : ParticleOpBase(dx, xlo, dy, ylo, dz, zlo) {}

  void apply_to_field(const SrcFieldType& src, DestFieldType& dest) const
  {
    const size_t np = dest.window_with_ghost().glob_npts();
    check(np);
    nebo_field_t s = src.descriptor();
    const short dev = (short)s.device;
    nebo_check(nebo_cell_to_particle(dest.field_values(dev), &s, on(px_, dev), on(py_, dev), on(pz_, dev),
                                     on(psize_, dev), (int64_t)np, &geom_, nullptr, dest.get_stream()));
    dest.mark_written();
  }
};

}  // namespace Particle
}  // namespace SpatialOps
