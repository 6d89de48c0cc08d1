// Host cost of the C++ front end on the time loop of examples/heat_equation.cpp:
// three assignments per step (the diffusive right-hand side, the Euler update,
// the boundary reset through masks), here on an n x n x n SVolField on the GPU
// (default 64^3) with the z terms added.  Each assignment builds its expression,
// serialises it and calls nebo_assign; repeated descriptors replay their launches.
// Prints one JSON line: wall time per step with the GPU waited for at the end,
// and the host time to submit the steps' assignments.
#include <spatialops/Nebo.h>
#include <spatialops/OperatorDatabase.h>

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace SpatialOps;
typedef SVolField FieldT;

int main(int argc, char** argv)
{
  const int n = argc > 1 ? std::atoi(argv[1]) : 64;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 200;
  if (n < 2 || steps < 1) {
    std::fprintf(stderr, "usage: heat_timing [n >= 2] [steps >= 1]\n");
    return 2;
  }
  const IntVec dim(n, n, n);
  const GhostData nghost(1);
  const BoundaryCellInfo bcInfo = BoundaryCellInfo::build<FieldT>(false, false, false);
  const MemoryWindow window(get_window_with_ghost(dim, nghost, bcInfo));
  FieldT phi(window, bcInfo, nghost, nullptr, InternalStorage, GPU_INDEX);
  FieldT rhs(window, bcInfo, nghost, nullptr, InternalStorage, GPU_INDEX);
  FieldT alpha(window, bcInfo, nghost, nullptr, InternalStorage, GPU_INDEX);
  const Grid grid(dim, {1.0, 1.0, 1.0});
  alpha <<= 1.0;
  rhs <<= 0.0;

  // the four side faces of the ghost layer, over every plane
  std::vector<IntVec> leftSet, rightSet, topSet, bottomSet;
  for (int z = -1; z <= n; ++z)
    for (int i = -1; i <= n; ++i) {
      leftSet.push_back(IntVec(-1, i, z));
      rightSet.push_back(IntVec(n, i, z));
      topSet.push_back(IntVec(i, -1, z));
      bottomSet.push_back(IntVec(i, n, z));
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
  const Ops::DivZ& divZ = *opDB.retrieve_operator<Ops::DivZ>();
  const Ops::GradZ& gradZ = *opDB.retrieve_operator<Ops::GradZ>();
  const Ops::InterpC2FZ& interpZ = *opDB.retrieve_operator<Ops::InterpC2FZ>();

  const double dx = grid.spacing<XDIR>();
  const double deltaT = 0.1 * dx * dx / field_max(alpha);
  phi <<= cond(left, 10.0)(right, 0.0)(5.0);

  auto step = [&] {
    rhs <<= divX(interpX(alpha) * gradX(phi)) + divY(interpY(alpha) * gradY(phi)) +
            divZ(interpZ(alpha) * gradZ(phi));
    phi <<= phi + deltaT * rhs;
    phi <<= cond(left, 10.0)(right, 0.0)(top || bottom, 5.0)(phi);
  };
  for (int i = 0; i < 5; ++i) step();   // the first calls compile and record
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  typedef std::chrono::steady_clock clock;
  const auto t0 = clock::now();
  for (int i = 0; i < steps; ++i) step();
  const auto t1 = clock::now();
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  const auto t2 = clock::now();
  const double submit = std::chrono::duration<double, std::micro>(t1 - t0).count() / steps;
  const double wall = std::chrono::duration<double, std::micro>(t2 - t0).count() / steps;
  const double mx = field_max_interior(phi), mn = field_min_interior(phi);
  std::printf("{\"n\": %d, \"steps\": %d, \"assignments_per_step\": 3, \"us_per_step\": %.3f, "
              "\"host_submit_us_per_step\": %.3f, \"interior_min_max\": [%.17g, %.17g]}\n",
              n, steps, wall, submit, mn, mx);
  return (mn >= 0.0 && mx <= 10.0) ? 0 : 1;
}
