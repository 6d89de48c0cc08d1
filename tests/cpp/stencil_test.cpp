// The reference's stencil tests, written against the C++ front end and run on
// the GPU:
//  * test/StencilTest.cpp: the fused chain
//      result <<= -divX(gradX(a)*interpX(b)) - divY(..) - divZ(..)
//    equals the unfused chain through face temporaries, exactly;
//  * spatialops/structured/stencil/test/test_ref_stencil.cpp: every basic
//    operator equals the hand loop s1*coefLo + s2*coefHi, exactly.
#include <spatialops/Nebo.h>
#include <spatialops/OperatorDatabase.h>

#include <cstdlib>
#include <vector>

#include "nebo_test_helper.h"

using namespace SpatialOps;
typedef SVolField Field;

template <typename F>
F make_field(const MemoryWindow& w, const BoundaryCellInfo& bc, const GhostData& g)
{
  return F(w, bc, g, nullptr, InternalStorage, CPU_INDEX);
}

int main(int argc, char** argv)
{
  TestHelper status;
  const int nx = argc > 1 ? std::atoi(argv[1]) : 23, ny = argc > 2 ? std::atoi(argv[2]) : 17,
            nz = argc > 3 ? std::atoi(argv[3]) : 11;
  const GhostData ghost(1);
  const BoundaryCellInfo bc = BoundaryCellInfo::build<Field>(false, false, false);
  const MemoryWindow window(get_window_with_ghost(IntVec(nx, ny, nz), ghost, bc));

  Field a = make_field<Field>(window, bc, ghost), b = make_field<Field>(window, bc, ghost);
  for (size_t k = 0; k < window.glob_dim(2); ++k)
    for (size_t j = 0; j < window.glob_dim(1); ++j)
      for (size_t i = 0; i < window.glob_dim(0); ++i) {
        a(i, j, k) = i + j * 2.0 + k * 4.0;
        b(i, j, k) = i + j * 3.0 + k * 5.0;
      }
  a.set_device_as_active(GPU_INDEX);
  b.set_device_as_active(GPU_INDEX);

  OperatorDatabase opDB;
  build_stencils(nx, ny, nz, 1.0, 1.5, 2.0, opDB);
  typedef BasicOpTypes<Field> Ops;
  const Ops::GradX& gradX = *opDB.retrieve_operator<Ops::GradX>();
  const Ops::GradY& gradY = *opDB.retrieve_operator<Ops::GradY>();
  const Ops::GradZ& gradZ = *opDB.retrieve_operator<Ops::GradZ>();
  const Ops::InterpC2FX& interpX = *opDB.retrieve_operator<Ops::InterpC2FX>();
  const Ops::InterpC2FY& interpY = *opDB.retrieve_operator<Ops::InterpC2FY>();
  const Ops::InterpC2FZ& interpZ = *opDB.retrieve_operator<Ops::InterpC2FZ>();
  const Ops::DivX& divX = *opDB.retrieve_operator<Ops::DivX>();
  const Ops::DivY& divY = *opDB.retrieve_operator<Ops::DivY>();
  const Ops::DivZ& divZ = *opDB.retrieve_operator<Ops::DivZ>();

  // unfused ("old") chain through temporaries
  FaceTypes<Field>::XFace fx(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX), fx2(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
  FaceTypes<Field>::YFace fy(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX), fy2(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
  FaceTypes<Field>::ZFace fz(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX), fz2(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
  Field tx(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX), ty(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX),
      tz(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
  Field sr(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX), cr(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
  gradX.apply_to_field(a, fx);
  interpX.apply_to_field(b, fx2);
  fx <<= fx * fx2;
  divX.apply_to_field(fx, tx);
  gradY.apply_to_field(a, fy);
  interpY.apply_to_field(b, fy2);
  fy <<= fy * fy2;
  divY.apply_to_field(fy, ty);
  gradZ.apply_to_field(a, fz);
  interpZ.apply_to_field(b, fz2);
  fz <<= fz * fz2;
  divZ.apply_to_field(fz, tz);
  sr <<= -tx - ty - tz;
  // fused ("new") chain
  cr <<= (-divX(gradX(a) * interpX(b)) - divY(gradY(a) * interpY(b)) - divZ(gradZ(a) * interpZ(b)));

  sr.set_device_as_active(CPU_INDEX);
  cr.set_device_as_active(CPU_INDEX);
  bool same = true;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) same = same && (cr(i, j, k) == sr(i, j, k));
  status(same, "StencilTest: fused chain == unfused chain (exact, interior)");

  // hand loops of the basic operators
  a.set_device_as_active(CPU_INDEX);
  std::vector<double> av(a.field_values(CPU_INDEX), a.field_values(CPU_INDEX) + window.glob_npts());
  a.set_device_as_active(GPU_INDEX);
  const IntVec G = window.glob_dim();
  auto A = [&](int i, int j, int k) { return av[i + G[0] * (j + G[1] * k)]; };
  {
    FaceTypes<Field>::XFace g(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
    g <<= gradX(a);
    g.set_device_as_active(CPU_INDEX);
    const std::vector<double>& c = gradX.coefs();
    bool ok = true;
    for (int k = 0; k < G[2]; ++k)
      for (int j = 0; j < G[1]; ++j)
        for (int i = 1; i < G[0]; ++i) ok = ok && (g(i, j, k) == A(i - 1, j, k) * c[0] + A(i, j, k) * c[1]);
    status(ok, "test_ref_stencil: GradX == hand loop");
  }
  {
    FaceTypes<Field>::YFace g(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
    g <<= interpY(a);
    g.set_device_as_active(CPU_INDEX);
    bool ok = true;
    for (int k = 0; k < G[2]; ++k)
      for (int j = 1; j < G[1]; ++j)
        for (int i = 0; i < G[0]; ++i) ok = ok && (g(i, j, k) == A(i, j - 1, k) * 0.5 + A(i, j, k) * 0.5);
    status(ok, "test_ref_stencil: InterpC2FY == hand loop");
  }
  {
    // DivZ of a face field whose values are the a-values
    FaceTypes<Field>::ZFace f(window, bc, ghost, nullptr, InternalStorage, CPU_INDEX);
    for (int k = 0; k < G[2]; ++k)
      for (int j = 0; j < G[1]; ++j)
        for (int i = 0; i < G[0]; ++i) f(i, j, k) = A(i, j, k);
    f.set_device_as_active(GPU_INDEX);
    Field d(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
    d <<= divZ(f);
    d.set_device_as_active(CPU_INDEX);
    const std::vector<double>& c = divZ.coefs();
    bool ok = true;
    for (int k = 0; k < G[2] - 1; ++k)
      for (int j = 0; j < G[1]; ++j)
        for (int i = 0; i < G[0]; ++i) ok = ok && (d(i, j, k) == A(i, j, k) * c[0] + A(i, j, k + 1) * c[1]);
    status(ok, "test_ref_stencil: DivZ == hand loop");
  }
  // the type system: a stencil chain over a 4-point interpolant and a box filter compiles and runs
  {
    typedef OperatorTypeBuilder<Interpolant, SVolField, XSurfYField>::type I4;
    typedef OperatorTypeBuilder<Filter, SVolField, SVolField>::type Box;
    XSurfYField q(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
    q <<= (*opDB.retrieve_operator<I4>())(a);
    Field fb(window, bc, ghost, nullptr, InternalStorage, GPU_INDEX);
    fb <<= (*opDB.retrieve_operator<Box>())(a);
    fb.set_device_as_active(CPU_INDEX);
    // the box average of a linear field is the field itself at interior points
    bool ok = true;
    for (int k = 1; k < G[2] - 1; ++k)
      for (int j = 1; j < G[1] - 1; ++j)
        for (int i = 1; i < G[0] - 1; ++i) ok = ok && std::fabs(fb(i, j, k) - A(i, j, k)) <= 1e-12 * (1 + A(i, j, k));
    status(ok, "box filter of a linear field reproduces it");
  }
  return status.report();
}
