"""The C++ front end (spatialops-threadfix_amd/cpp/include): the reference's API on
the engine.  CPU: it compiles and rejects mixed field types at compile time, as
the reference's NeboFieldCheck does.  GPU: the reference-style test programs
(tests/cpp/*.cpp, built by __graft_entry__.build()) pass."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "spatialops-threadfix_amd", "cpp")
BIN = os.path.join(CPP, "bin")
INC = ["-I", os.path.join(CPP, "include"), "-I", os.path.join(ROOT, "include")]


def _compile(src, tmp_path):
    f = tmp_path / "t.cpp"
    f.write_text(src)
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", *INC, str(f)],
                          capture_output=True, text=True)


GOOD = r"""
#include <spatialops/structured/FVStaggered.h>
using namespace SpatialOps;
void f(SVolField& r, const SVolField& k, const SVolField& phi, const OperatorDatabase& db) {
  typedef BasicOpTypes<SVolField> Ops;
  const Ops::DivX& dx = *db.retrieve_operator<Ops::DivX>();
  const Ops::GradX& gx = *db.retrieve_operator<Ops::GradX>();
  const Ops::InterpC2FX& ix = *db.retrieve_operator<Ops::InterpC2FX>();
  r <<= dx(ix(k) * gx(phi)) + 2.0 * phi;
  r <<= cond(phi > 0.0, phi * phi)(-phi);
  interior_assign(r, sqrt(abs(phi)) + max(phi, k));
  double m = field_norm_interior(phi) + field_max(phi * k);
  (void)m;
}
void g(SVolField& phi, const SSurfXField& gamma, const SpatialMask<SSurfXField>& mask,
       const OperatorDatabase& db) {
  typedef BasicOpTypes<SVolField>::GradX OpT;
  NeboBoundaryConditionBuilder<OpT> bc(*db.retrieve_operator<OpT>());
  bc(mask, phi, gamma, true);
  bc(mask, phi, 2.0, false);
  NeboMaskShiftBuilder<OpT, SSurfXField, SVolField> shift;
  SVolField out(phi);
  out <<= cond(shift.minus(mask), 1.0)(0.0);
  typedef OperatorTypeBuilder<GradientY, SVolField, SVolField>::type FD;
  NeboBoundaryConditionBuilder<FD> fd(*db.retrieve_operator<FD>());
  SpatialMask<SVolField> m2 = SpatialMask<SVolField>::build(phi, {IntVec(0, 0, 0)});
  fd(m2, phi, phi * 2.0, true);
}
"""

BAD = r"""
#include <spatialops/structured/FVStaggered.h>
using namespace SpatialOps;
void f(SVolField& r, const SVolField& a, const SSurfXField& b) { r <<= a + b; }
"""


SIGNATURES = r"""
#include <spatialops/structured/FVStaggered.h>
#include <type_traits>
using namespace SpatialOps;
// user code written against the reference's signatures (NeboAssignment.h:29-133,
// NeboBasic.h:193-243): it names NeboExpression<RhsType, FieldType> and
// NeboBooleanExpression<RhsType, FieldType> and binds FieldType const& results
template <typename RhsType, typename FieldType>
FieldType const& assign_expr(FieldType& lhs, const NeboExpression<RhsType, FieldType>& rhs)
{
  return lhs <<= rhs;
}
template <typename RhsType, typename FieldType>
FieldType const& select(FieldType& out, const NeboBooleanExpression<RhsType, FieldType>& test)
{
  return out <<= cond(test, 1.0)(0.0);
}
void f(SVolField& r, const SVolField& a, const SVolField& b, const SpatialMask<SVolField>& m)
{
  SVolField const& x = (r <<= a + b);
  SVolField const& y = assign_expr(r, a * 2.0);
  SVolField const& z = interior_assign(r, a - b);
  SVolField const& w = masked_assign(m, r, 3.0);
  SVolField const& v = select(r, a > b);
  auto e = a * b + 1.0;
  static_assert(std::is_same<decltype(e)::field_type, SVolField>::value, "field_type");
  typedef decltype(e)::Expression RhsT;
  NeboExpression<RhsT, SVolField> copy(e.expr());
  r <<= copy;
  r <<= 2.0;
  r <<= a;
  (void)x; (void)y; (void)z; (void)w; (void)v;
}
"""


def test_reference_signatures_compile(tmp_path):
    """Reference-style user code that names the two-parameter expression types and
    binds the FieldType const& that <<=, interior_assign and masked_assign return."""
    r = _compile(SIGNATURES, tmp_path)
    assert r.returncode == 0, r.stderr


def test_front_end_compiles(tmp_path):
    r = _compile(GOOD, tmp_path)
    assert r.returncode == 0, r.stderr


def test_mixed_field_types_do_not_compile(tmp_path):
    r = _compile(BAD, tmp_path)
    assert r.returncode != 0
    assert "same field type" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("prog", ["stencil_test", "mask_test", "heat_equation", "reduction_test", "bc_test",
                                  "store_test", "particle_test"])
def test_cpp_program(prog):
    exe = os.path.join(BIN, prog)
    assert os.path.exists(exe), f"{exe} not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "ALL TESTS PASSED" in r.stdout
