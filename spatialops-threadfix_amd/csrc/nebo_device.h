// nebo_device.h -- device helpers shared by the hand-written stencil kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace nebo {

// Hardware dispatches workgroup b to XCD b % 8.  Map b to a logical tile so that
// each XCD receives one contiguous run of logical tiles (neighbouring tiles share
// halo rows and planes, and each XCD has its own L2).
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, k = b >> 3;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Tile grid over the n[0] x n[1] x n[2] interior of a periodic block: tiles of
// 64 x TY columns marching ZC planes.
struct Box3 {
  int n[3];
  int nx, ny, nz;  // tile counts
  int zc;          // planes per tile
};

inline Box3 make_box(const int n[3], int tx, int ty, int zc)
{
  Box3 B;
  for (int d = 0; d < 3; ++d) B.n[d] = n[d];
  B.zc = zc;
  B.nx = (n[0] + tx - 1) / tx;
  B.ny = (n[1] + ty - 1) / ty;
  B.nz = (n[2] + zc - 1) / zc;
  return B;
}

// interp(k)*grad(phi) on one face: (k_lo*i0 + k_hi*i1) * (p_lo*g0 + p_hi*g1)
__device__ __forceinline__ double face_flux(double klo, double khi, double plo, double phi,
                                            double i0, double i1, double g0, double g1)
{
  const double ki = klo * i0 + khi * i1;
  const double gr = plo * g0 + phi * g1;
  return ki * gr;
}

}  // namespace nebo
