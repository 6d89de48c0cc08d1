// nebo_device.h -- device helpers shared by the hand-written stencil kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "../../include/nebo_hip.h"

namespace nebo {

constexpr int kLanes = 64;   // wavefront width of gfx9 (wave64)

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

// ---- column-schedule helpers (shared by the stencil and transport kernels) ----

struct Tiles {
  int lo[3], hi[3];
  int nx, ny, nz;  // tile counts
};

inline Tiles make_tiles(const nebo_limits_t& L, int tx, int ty, int zc)
{
  Tiles t;
  for (int d = 0; d < 3; ++d) { t.lo[d] = L.lo[d]; t.hi[d] = L.hi[d]; }
  t.nx = (L.hi[0] - L.lo[0] + tx - 1) / tx;
  t.ny = (L.hi[1] - L.lo[1] + ty - 1) / ty;
  t.nz = (L.hi[2] - L.lo[2] + zc - 1) / zc;
  return t;
}


// Lane l receives lane l-1's value (DPP wave_shr:1); lane 0 receives 0 (every
// caller replaces lane 0's value).
__device__ __forceinline__ double from_lower_lane(double v)
{
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);
}

// Lane l receives lane l+1's value (DPP wave_shl:1); lane 63 receives 0.
__device__ __forceinline__ double from_upper_lane(double v)
{
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);
}


// Lane j's value of v (v_readlane; j is uniform), as seen by every lane.
// Lane j's value of v, for every lane.  Take it outside any branch or select on
// the lane: inside one, it runs with only some lanes active, and a value the
// compiler keeps in other registers (a spill) is brought back for those lanes
// only, so lane j may read a register nobody refilled.
__device__ __forceinline__ double lane_value(double v, int j)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
  return __hiloint2double(hi, lo);
}


// Loads of one column-schedule lane within a plane, clamped into [lo-1, hi] on
// every axis: rows y0-1 .. y0+RY of its column, and one edge load: lanes 0..RY-1
// take the left neighbours (x = xw0-1) of rows 0..RY-1, lanes 32..32+RY-1 the
// right neighbours (x = xw0+64), other lanes re-read their row 0.  Offsets are in
// bytes from the plane's base, the point (lo[0]-1, lo[1]-1) of the plane (see
// plane_base), so they are non-negative 32-bit values whatever the limits (an
// assignment that includes ghost cells starts below the frame origin) and every
// load is base (uniform) + offset; for an active lane the offset of row j+1 is
// also the store offset of row j.
template <int RY>
struct ColumnGeom {
  uint32_t row[RY + 2];
  uint32_t edge;
  uint32_t inner;   // lanes 0..RY-1: (xw0, row); lanes 32..32+RY-1: (xw0+63, row); else row 0
};

template <int RY>
__device__ __forceinline__ ColumnGeom<RY> column_geom(const Tiles& T, int X, int xw0, int y0, int lane)
{
  ColumnGeom<RY> g;
  const int ax = T.lo[0] - 1, ay = T.lo[1] - 1;
  auto off = [X, ax, ay](int x, int y) { return (uint32_t)((x - ax) + X * (y - ay)) * 8u; };
  const int xc = min(xw0 + lane, T.hi[0]);
#pragma unroll
  for (int r = 0; r < RY + 2; ++r) g.row[r] = off(xc, min(y0 - 1 + r, T.hi[1]));
  // (a wave wholly beyond hi, in a partial last tile, has xw0 - 1 > hi: clamp too)
  if (lane < RY) g.edge = off(min(xw0 - 1, T.hi[0]), min(y0 + lane, T.hi[1]));
  else if (lane >= 32 && lane < 32 + RY) g.edge = off(min(xw0 + kLanes, T.hi[0]), min(y0 + lane - 32, T.hi[1]));
  else g.edge = g.row[1];
  if (lane < RY) g.inner = off(min(xw0, T.hi[0]), min(y0 + lane, T.hi[1]));
  else if (lane >= 32 && lane < 32 + RY) g.inner = off(min(xw0 + kLanes - 1, T.hi[0]), min(y0 + lane - 32, T.hi[1]));
  else g.inner = g.row[1];
  return g;
}

// Plane z's base: the point (lo[0]-1, lo[1]-1, z) of the frame, the lowest
// point of the plane any stencil of the tile reads (uniform across the wave).
template <typename V>
__device__ __forceinline__ V* plane_base(V* origin, const Tiles& T, int X, int XY, int z)
{
  return origin + ((long)XY * z + (T.lo[0] - 1) + (long)X * (T.lo[1] - 1));
}

// Raw buffer resource over one plane (base from plane_base): loads and stores
// take the lane's 32-bit byte offset, so no per-lane address arithmetic is left
// in the loop.  The range, (XY + 2X + 2) doubles, covers every offset the column
// geometry forms (the points [lo-1, hi] of a box inside the block lie less than
// XY points past the base; the hardware returns 0 for a load beyond the range
// instead of faulting; the launcher keeps it below 2^31 bytes).  Word 3 is the
// gfx9 raw buffer format.
using Rsrc = __amdgpu_buffer_rsrc_t;
using U2 = __attribute__((__vector_size__(2 * sizeof(unsigned int)))) unsigned int;

__device__ __forceinline__ Rsrc plane_rsrc(const double* base, int X, int XY)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, (XY + 2 * X + 2) * 8, 0x00020000);
}

__device__ __forceinline__ double load_at(Rsrc rs, uint32_t off)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
}

// nontemporal store (cache-policy bit 1: nt)
__device__ __forceinline__ void store_nt_at(Rsrc rs, uint32_t off, double v)
{
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), rs, off, 0, 2);
}

// store with cache policy SP: 2 nontemporal, 0 write-back (the line stays in L2
// until evicted, so neighbouring waves' parts of one line are merged there)
template <int SP>
__device__ __forceinline__ void store_at(Rsrc rs, uint32_t off, double v)
{
  static_assert(SP == 0 || SP == 2, "store policy");
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), rs, off, 0, SP);
}

// Two adjacent doubles moved by one 16-byte buffer access (the address needs only
// 8-byte alignment: compute queues run with unaligned access enabled).
typedef double D2 __attribute__((ext_vector_type(2)));
typedef unsigned U4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ D2 load2_at(Rsrc rs, uint32_t off)
{
  return __builtin_bit_cast(D2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
}

// nontemporal 16-byte store
__device__ __forceinline__ void store2_nt_at(Rsrc rs, uint32_t off, D2 v)
{
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, v), rs, off, 0, 2);
}

// Offsets of one pair-column lane (two adjacent columns x, x + 1 with
// x = xw0 + 2 lane) within a plane, from the plane base of plane_base: rows
// y0-1 .. y0+RY, and one edge load: lanes 0..RY-1 take the left neighbours
// (x = xw0 - 1) of rows 0..RY-1, lanes 32..32+RY-1 the right neighbours
// (x = xw0 + 128), other lanes re-read their row 0.  The box width is a multiple
// of the 128-column wave, so every lane's columns lie inside it; rows past the
// box are clamped to the last readable row.
template <int RY>
struct PairGeom {
  uint32_t row[RY + 2];
  uint32_t edge;
};

template <int RY>
__device__ __forceinline__ PairGeom<RY> pair_geom(const Tiles& T, int X, int xw0, int y0, int lane)
{
  PairGeom<RY> g;
  const int ax = T.lo[0] - 1, ay = T.lo[1] - 1;
  auto off = [X, ax, ay](int x, int y) { return (uint32_t)((x - ax) + X * (y - ay)) * 8u; };
  const int x = xw0 + 2 * lane;
#pragma unroll
  for (int r = 0; r < RY + 2; ++r) g.row[r] = off(x, min(y0 - 1 + r, T.hi[1]));
  if (lane < RY) g.edge = off(xw0 - 1, min(y0 + lane, T.hi[1]));
  else if (lane >= 32 && lane < 32 + RY) g.edge = off(xw0 + 2 * kLanes, min(y0 + lane - 32, T.hi[1]));
  else g.edge = g.row[1];
  return g;
}

// Run f(integral_constant<int, I>) for I = 0, 1, ... in order (compile-time steps).
template <typename F, int... I>
__device__ __forceinline__ void unroll_steps(F&& f, std::integer_sequence<int, I...>)
{
  (f(std::integral_constant<int, I>{}), ...);
}

}  // namespace nebo
