// spatialops/particles/test/TestParticleOperators.cpp and
// spatialops/pointfield/test/testPointField.cpp against the C++ front end on the
// GPU: one particle of size 5 at x = 3.25 over a 10-cell line with 3 ghost cells
// (c2p of 10*x gives 32.5, p2c of 20 gives 1,4,4,4,4,3,0,0,0,0), and the point
// field's index gather / scatter.  (The reference's point test sets f[2] and f[4]
// through its old point boundary-condition tool; here they are set directly.)
#include <spatialops/structured/FVStaggered.h>
#include <spatialops/particles/ParticleOperators.h>
#include <spatialops/pointfield/PointOperators.h>

#include "nebo_test_helper.h"

using namespace SpatialOps;
typedef SVolField CellField;

int main()
{
  TestHelper status;
  {
    const IntVec dim(10, 1, 1);
    const Grid grid(dim, {10.0, 1.0, 1.0});
    const GhostData cg(3);
    const BoundaryCellInfo cbc = BoundaryCellInfo::build<CellField>();
    const MemoryWindow mw = get_window_with_ghost(dim, cg, BoundaryCellInfo::build<CellField>(false, false, false));
    CellField xcoord(mw, cbc, cg, nullptr);
    CellField f(mw, cbc, cg, nullptr);
    CellField ctmp(mw, cbc, cg, nullptr);

    const GhostData pg(0);
    const BoundaryCellInfo pbc = BoundaryCellInfo::build<Particle::ParticleField>();
    const MemoryWindow pmw(IntVec(1, 1, 1));
    Particle::ParticleField pCoord(pmw, pbc, pg, nullptr), pSize(pmw, pbc, pg, nullptr);
    Particle::ParticleField pfield(pmw, pbc, pg, nullptr), ptmp(pmw, pbc, pg, nullptr);

    grid.set_coord<XDIR>(xcoord);
    pCoord[0] = 3.25;
    pSize[0] = 5;
    pfield[0] = 20;
    const double dx = xcoord[1] - xcoord[0], xlo = xcoord[cg.get_minus(0)];

    for (CellField* c : {&xcoord, &f, &ctmp}) c->set_device_as_active(GPU_INDEX);
    for (Particle::ParticleField* p : {&pCoord, &pSize, &pfield, &ptmp}) p->set_device_as_active(GPU_INDEX);
    f <<= 10 * xcoord;

    Particle::CellToParticle<CellField> c2p(dx, xlo);
    c2p.set_coordinate_information(&pCoord, nullptr, nullptr, &pSize);
    Particle::ParticleToCell<CellField> p2c(dx, xlo);
    p2c.set_coordinate_information(&pCoord, nullptr, nullptr, &pSize);
    c2p.apply_to_field(f, ptmp);
    p2c.apply_to_field(pfield, ctmp);

    ptmp.set_device_as_active(CPU_INDEX);
    ctmp.set_device_as_active(CPU_INDEX);
    status(ptmp[0] == 32.5, "c2p");
    const size_t ishift = cg.get_minus(0);
    const double want[10] = {1, 4, 4, 4, 4, 3, 0, 0, 0, 0};
    for (int i = 0; i < 10; ++i) status(ctmp[ishift + i] == want[i], "p2c [" + std::to_string(i) + "]");
  }
  {
    typedef Point::PointField FieldT;
    const size_t npts = 10;
    const GhostData ghost(0);
    const BoundaryCellInfo bc = BoundaryCellInfo::build<FieldT>();
    const MemoryWindow mw(IntVec(npts, 1, 1));
    FieldT f(mw, bc, ghost, nullptr);
    double x = 0.1;
    for (FieldT::iterator ifld = f.begin(); ifld != f.end(); ++ifld, x += 1.0) *ifld = x;
    FieldT::iterator i2 = f.interior_begin();
    bool same = true;
    for (FieldT::iterator i = f.begin(); i != f.end(); ++i, ++i2) same = same && *i == *i2 && &*i == &*i2;
    status(same, "iterators: values and addresses");
    f[2] = 1.234;
    f[4] = 3.456;

    std::vector<size_t> ix = {4, 2};
    Point::FieldToPoint<FieldT> ftp(ix);
    Point::PointToField<FieldT> ptf(ix);
    const MemoryWindow mw2(IntVec(2, 1, 1));
    FieldT f2(mw2, bc, ghost, nullptr);
    f.set_device_as_active(GPU_INDEX);
    f2.set_device_as_active(GPU_INDEX);
    ftp.apply_to_field(f, f2);
    f2.set_device_as_active(CPU_INDEX);
    status(f2[0] == 3.456, "Field2Point Interp (1)");
    status(f2[1] == 1.234, "Field2Point Interp (2)");
    f2[0] = -1.234;
    f2[1] = -3.456;
    f2.set_device_as_active(GPU_INDEX);
    ptf.apply_to_field(f2, f);
    f.set_device_as_active(CPU_INDEX);
    status(f[2] == -3.456, "Point2Field Interp (1)");
    status(f[4] == -1.234, "Point2Field Interp (2)");
  }
  return status.report();
}
