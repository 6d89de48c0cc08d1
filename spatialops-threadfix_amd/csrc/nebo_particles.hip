// nebo_particles.hip -- particle <-> cell interpolation and point fields.
//
// CellToParticle is a gather: one thread per particle walks the cells its cube
// overlaps (k, then j, then i, as ParticleOperators.h:339-388 does) and sums
// src(i,j,k) * fraction into a register.  ParticleToCell is a scatter whose
// per-cell sums must come out in particle order to match the reference's serial
// loop bit for bit (ParticleOperators.h:184-270), so atomics are out.  Instead:
//   1. count the cells each particle overlaps, exclusive scan -> entry offsets;
//   2. each particle writes its entries in its own loop order: value
//      src[p] * fraction and key cell at entry e, so entries stand in the serial
//      loop's order;
//   3. stable radix sort of (cell, value) pairs on the cell's bits alone:
//      entries of one cell become contiguous and keep their serial order;
//   4. the first entry of each cell sums the cell's run onto the cell's value,
//      reading the sorted values in sequence.
//
// Point fields (pointfield/PointOperators.h:40-160) are an index gather and an
// index scatter; repeated scatter indices keep the last point, resolved on the
// host when the index list is uploaded.

#include "nebo_internal.h"

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace nebo {
namespace {

constexpr int kBlock = 256;

// One cell field as the particle loops see it: cell (i,j,k) of the window
// (0-based, ghosts included) lives at flat index (i+o0) + X*((j+o1) + Y*(k+o2)),
// an axis of glob extent 1 always at 0 (MemoryWindow::flat_index,
// MemoryWindow.h:121-131).
struct CellGrid {
  int nmax[3];       // window extent
  int off[3];        // window offset
  int glob[3];       // glob extents
  double h[3];       // spacing
  double loface[3];  // coordinate of the (-) face of window cell 0
};

CellGrid make_grid(const nebo_field_t* f, const nebo_particle_geom_t& g)
{
  CellGrid c;
  const double d[3] = {g.dx, g.dy, g.dz}, lo[3] = {g.xlo, g.ylo, g.zlo};
  for (int a = 0; a < 3; ++a) {
    c.nmax[a] = f->extent[a];
    c.off[a] = f->offset[a];
    // the constructor's defaults: spacing 1 and origin 0 where the spacing is <= 0
    c.h[a] = d[a] <= 0 ? 1.0 : d[a];
    const double o = d[a] <= 0 ? 0.0 : lo[a];
    // xlo - dx*ngm - dx/2, evaluated left to right
    c.loface[a] = o - c.h[a] * f->ghost_minus[a] - c.h[a] / 2;
  }
  for (int a = 0; a < 3; ++a) c.glob[a] = f->glob[a];
  return c;
}

// The box of cells a particle overlaps and what the fractions need.
struct PBox {
  long lo[3], hi[3];
  double plo[3], phi[3];
  double pvol;
  bool has[3];
  bool skip;
};

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }   // std::min
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }   // std::max

__device__ __forceinline__ PBox particle_box(const CellGrid& g, const double* const coord[3],
                                             const double* psize, long p)
{
  PBox b;
  const double size = psize[p];
  const double rp = size * 0.5;
  b.skip = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b.has[a] = coord[a] != nullptr;
    b.plo[a] = b.has[a] ? coord[a][p] - rp : 0;
    b.phi[a] = b.has[a] ? b.plo[a] + size : 0;
    // size_t conversion of the low index: exact for t > -1 (truncation); at
    // t <= -1 (or NaN) the reference's conversion wraps and its loop is empty
    const double t = (b.plo[a] - g.loface[a]) / g.h[a];
    if (!(t > -1.0)) b.skip = true;
    const double tc = t > 4.0e18 ? 4.0e18 : (t > 0.0 ? t : 0.0);
    b.lo[a] = (long)tc;
    if (b.has[a]) {
      // std::min(nmax, int(u) + 1), with u clamped where int(u) would overflow
      double u = (b.phi[a] - g.loface[a]) / g.h[a];
      u = u > 2.0e9 ? 2.0e9 : (u < -2.0e9 ? -2.0e9 : u);
      b.hi[a] = (long)u + 1;
    } else {
      b.hi[a] = b.lo[a] + 1;
    }
    // the reference bounds only a resolved axis by the window; bounding every
    // axis keeps each access inside the window (the same cells where the
    // reference's are in bounds)
    if (b.hi[a] > g.nmax[a]) b.hi[a] = g.nmax[a];
  }
  b.pvol = (b.has[0] ? 2 * rp : 1) * (b.has[1] ? 2 * rp : 1) * (b.has[2] ? 2 * rp : 1);
  return b;
}

__device__ __forceinline__ long box_cells(const PBox& b)
{
  if (b.skip) return 0;
  long n = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) n *= b.hi[a] > b.lo[a] ? b.hi[a] - b.lo[a] : 0;
  return n;
}

// Visit the cells of the box in the reference's order with (flat index, fraction).
template <typename F>
__device__ __forceinline__ void for_cells(const CellGrid& g, const PBox& b, F&& f)
{
  for (long k = b.lo[2]; k < b.hi[2]; ++k) {
    const double zcm = g.loface[2] + (double)k * g.h[2];
    const double zcp = zcm + g.h[2];
    const double zcont = b.has[2] ? dmin(zcp, b.phi[2]) - dmax(zcm, b.plo[2]) : 1;
    for (long j = b.lo[1]; j < b.hi[1]; ++j) {
      const double ycm = g.loface[1] + (double)j * g.h[1];
      const double ycp = ycm + g.h[1];
      const double ycont = b.has[1] ? dmin(ycp, b.phi[1]) - dmax(ycm, b.plo[1]) : 1;
      for (long i = b.lo[0]; i < b.hi[0]; ++i) {
        const double xcm = g.loface[0] + (double)i * g.h[0];
        const double xcp = xcm + g.h[0];
        const double xcont = b.has[0] ? dmin(xcp, b.phi[0]) - dmax(xcm, b.plo[0]) : 1;
        const double contribution = xcont * ycont * zcont / b.pvol;
        const long gi = g.glob[0] > 1 ? i + g.off[0] : 0;
        const long gj = g.glob[1] > 1 ? j + g.off[1] : 0;
        const long gk = g.glob[2] > 1 ? k + g.off[2] : 0;
        f(gi + (long)g.glob[0] * (gj + (long)g.glob[1] * gk), contribution);
      }
    }
  }
}

struct Coords {
  const double* c[3];
};

__global__ void cell_to_particle_kernel(double* __restrict__ dest, const double* __restrict__ src, CellGrid g,
                                        Coords pc, const double* __restrict__ psize, long np,
                                        int* __restrict__ skipped)
{
  const long p = (long)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const PBox b = particle_box(g, pc.c, psize, p);
  double acc = 0.0;
  if (!b.skip) for_cells(g, b, [&](long cell, double c) { acc += src[cell] * c; });
  else if (skipped) atomicAdd(skipped, 1);
  dest[p] = acc;
}

__global__ void count_kernel(long* __restrict__ counts, CellGrid g, Coords pc, const double* __restrict__ psize,
                             long np, int* __restrict__ skipped)
{
  const long p = (long)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const PBox b = particle_box(g, pc.c, psize, p);
  if (b.skip && skipped) atomicAdd(skipped, 1);
  counts[p] = box_cells(b);
}

// Entry e (the serial loop's e-th contribution: particles in order, each one's
// cells in k, j, i order) holds its cell and its value.
__global__ void fill_kernel(uint32_t* __restrict__ keys, double* __restrict__ vals,
                            const long* __restrict__ offsets, const double* __restrict__ src, CellGrid g,
                            Coords pc, const double* __restrict__ psize, long np)
{
  const long p = (long)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const PBox b = particle_box(g, pc.c, psize, p);
  if (b.skip) return;
  long e = offsets[p];
  const double s = src[p];
  for_cells(g, b, [&](long cell, double c) {
    vals[e] = s * c;
    keys[e] = (uint32_t)cell;
    ++e;
  });
}

// After a stable sort by cell, each cell's entries are contiguous and in serial
// order; the first entry of each run adds the run, in order, onto the cell.
__global__ void run_sum_kernel(double* __restrict__ dest, const uint32_t* __restrict__ keys,
                               const double* __restrict__ vals, long total)
{
  const long s = (long)blockIdx.x * kBlock + threadIdx.x;
  if (s >= total) return;
  const uint32_t cell = keys[s];
  if (s > 0 && keys[s - 1] == cell) return;
  double acc = dest[cell];
  // The run is read U entries at a time, all loads issued before the adds, which
  // still go one by one in entry order up to the first entry of another cell
  // (cells are below 2^31, so the padding key 0xffffffff ends a run).
  constexpr int U = 8;
  for (long t = s;; t += U) {
    uint32_t k[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      k[u] = t + u < total ? keys[t + u] : 0xffffffffu;
      v[u] = t + u < total ? vals[t + u] : 0.0;
    }
    bool end = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      end = end || k[u] != cell;
      if (!end) acc += v[u];
    }
    if (end) break;
  }
  dest[cell] = acc;
}

int bits_for(uint64_t n)   // bits to hold 0..n-1
{
  int b = 0;
  while (b < 64 && (1ull << b) < n) ++b;
  return b;
}

int check_particle_call(const nebo_field_t* f, const double* psize, int64_t np, const nebo_particle_geom_t* geom,
                        const char* ctx)
{
  if (!f || !geom || (np > 0 && !psize)) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  if (np < 0) return set_error(NEBO_ERR_INVALID, ctx, "negative particle count");
  if (f->device < 0) return set_error(NEBO_ERR_DEVICE, ctx, "the cell field must be resident on a GPU");
  if ((long)f->glob[0] * f->glob[1] * f->glob[2] >= (1L << 31))
    return set_error(NEBO_ERR_TOO_LARGE, ctx, "field has 2^31 or more points");
  return NEBO_OK;
}

}  // namespace

struct PointMap {
  int device;
  int64_t n;              // points
  int64_t max_index;
  int64_t* gather;        // n indices (FieldToPoint)
  int64_t* scatter_src;   // m point positions (PointToField, last occurrence of each index)
  int64_t* scatter_dst;   // m field indices
  int64_t m;
};

}  // namespace nebo

struct nebo_point_map : nebo::PointMap {};

using namespace nebo;

extern "C" int nebo_cell_to_particle(double* dest, const nebo_field_t* src, const double* px, const double* py,
                                     const double* pz, const double* psize, int64_t np,
                                     const nebo_particle_geom_t* geom, int32_t* skipped, void* stream)
{
  static const char* ctx = "Nebo Particle Interpolation";
  clear_error();
  int rc = check_particle_call(src, psize, np, geom, ctx);
  if (rc) return rc;
  if (np == 0) return NEBO_OK;
  if (!dest) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  const CellGrid g = make_grid(src, *geom);
  const Coords pc{{px, py, pz}};
  const long blocks = (np + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(cell_to_particle_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                     dest, static_cast<const double*>(src->values), g, pc, psize, (long)np, skipped);
  return check_hip(hipGetLastError(), ctx);
}

extern "C" int nebo_particle_to_cell(const nebo_field_t* dest, const double* src, const double* px,
                                     const double* py, const double* pz, const double* psize, int64_t np,
                                     const nebo_particle_geom_t* geom, int32_t* skipped, void* stream)
{
  static const char* ctx = "Nebo Particle Interpolation";
  clear_error();
  int rc = check_particle_call(dest, psize, np, geom, ctx);
  if (rc) return rc;
  if (np > 0 && !src) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  // dest <<= 0.0 (the reference's first step; same limits as any scalar assignment)
  {
    nebo_node_t zero;
    std::memset(&zero, 0, sizeof zero);
    zero.op = NEBO_SCALAR;
    zero.value = 0.0;
    const nebo_expr_t e{1, &zero, 0, nullptr, 0, nullptr, 0, nullptr};
    if ((rc = nebo_assign(dest, &e, 1, stream))) return rc;
    clear_error();
  }
  if (np == 0) return NEBO_OK;
  hipStream_t s = (hipStream_t)stream;
  const CellGrid g = make_grid(dest, *geom);
  const Coords pc{{px, py, pz}};
  const unsigned blocks = (unsigned)((np + kBlock - 1) / kBlock);
  long* counts = nullptr;
  long* offsets = nullptr;
  if ((rc = check_hip(hipMallocAsync((void**)&counts, sizeof(long) * 2 * np, s), ctx))) return rc;
  offsets = counts + np;
  hipLaunchKernelGGL(count_kernel, dim3(blocks), dim3(kBlock), 0, s, counts, g, pc, psize, (long)np, skipped);
  size_t tmp_bytes = 0;
  void* tmp = nullptr;
  rc = check_hip(rocprim::exclusive_scan(nullptr, tmp_bytes, counts, offsets, 0L, (size_t)np,
                                         rocprim::plus<long>(), s),
                 ctx);
  if (!rc) rc = check_hip(hipMallocAsync(&tmp, tmp_bytes, s), ctx);
  if (!rc)
    rc = check_hip(rocprim::exclusive_scan(tmp, tmp_bytes, counts, offsets, 0L, (size_t)np,
                                           rocprim::plus<long>(), s),
                   ctx);
  long last[2] = {0, 0};
  if (!rc) rc = check_hip(hipMemcpyAsync(&last[0], offsets + np - 1, sizeof(long), hipMemcpyDeviceToHost, s), ctx);
  if (!rc) rc = check_hip(hipMemcpyAsync(&last[1], counts + np - 1, sizeof(long), hipMemcpyDeviceToHost, s), ctx);
  if (!rc) rc = check_hip(hipStreamSynchronize(s), ctx);
  if (tmp) (void)hipFreeAsync(tmp, s);
  tmp = nullptr;
  const long total = last[0] + last[1];
  if (rc || total == 0) {
    (void)hipFreeAsync(counts, s);
    return rc;
  }
  // cells fit 32-bit keys (the field has fewer than 2^31 points, checked above)
  const long npts = (long)dest->glob[0] * dest->glob[1] * dest->glob[2];
  const int cbits = bits_for((uint64_t)npts);
  uint32_t* keys = nullptr;
  double* vals = nullptr;
  rc = check_hip(hipMallocAsync((void**)&keys, sizeof(uint32_t) * 2 * total, s), ctx);
  if (!rc) rc = check_hip(hipMallocAsync((void**)&vals, sizeof(double) * 2 * total, s), ctx);
  if (!rc) {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(kBlock), 0, s, keys, vals, offsets, src, g, pc, psize,
                       (long)np);
    rc = check_hip(hipGetLastError(), ctx);
  }
  // stable LSD radix sort on the cell alone: within a cell the entries keep
  // their serial order, so the run sums add in the serial loop's order
  uint32_t* skeys = keys ? keys + total : nullptr;
  double* svals = vals ? vals + total : nullptr;
  if (!rc)
    rc = check_hip(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, skeys, vals, svals, (size_t)total, 0,
                                             (unsigned)cbits, s),
                   ctx);
  if (!rc) rc = check_hip(hipMallocAsync(&tmp, tmp_bytes, s), ctx);
  if (!rc)
    rc = check_hip(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, skeys, vals, svals, (size_t)total, 0,
                                             (unsigned)cbits, s),
                   ctx);
  if (!rc) {
    hipLaunchKernelGGL(run_sum_kernel, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       static_cast<double*>(dest->values), skeys, svals, total);
    rc = check_hip(hipGetLastError(), ctx);
  }
  if (tmp) (void)hipFreeAsync(tmp, s);
  if (vals) (void)hipFreeAsync(vals, s);
  if (keys) (void)hipFreeAsync(keys, s);
  (void)hipFreeAsync(counts, s);
  return rc;
}

namespace {

__global__ void gather_kernel(double* __restrict__ dest, const double* __restrict__ src,
                              const int64_t* __restrict__ idx, long n)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dest[i] = src[idx[i]];
}

__global__ void scatter_kernel(double* __restrict__ dest, const double* __restrict__ src,
                               const int64_t* __restrict__ from, const int64_t* __restrict__ to, long m)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < m) dest[to[i]] = src[from[i]];
}

int check_map_field(const nebo_point_map* map, const nebo_field_t* f, const char* ctx)
{
  if (!map || !f) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  if (f->device != map->device) return set_error(NEBO_ERR_DEVICE, ctx, "field and index list are on different devices");
  const long npts = (long)f->glob[0] * f->glob[1] * f->glob[2];
  if (map->n > 0 && map->max_index >= npts)
    return set_error(NEBO_ERR_INVALID, ctx, "point index beyond the field's block");
  return NEBO_OK;
}

}  // namespace

extern "C" int nebo_point_map_create(const int64_t* indices, int64_t n, int device, nebo_point_map_t** out)
{
  static const char* ctx = "Nebo Point Field";
  clear_error();
  if (!out || (n > 0 && !indices)) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  if (n < 0) return set_error(NEBO_ERR_INVALID, ctx, "negative point count");
  *out = nullptr;
  std::vector<int64_t> from, to;
  std::unordered_map<int64_t, size_t> slot;   // field index -> position in from/to
  int64_t mx = -1;
  for (int64_t i = 0; i < n; ++i) {
    if (indices[i] < 0) return set_error(NEBO_ERR_INVALID, ctx, "negative point index");
    mx = indices[i] > mx ? indices[i] : mx;
    auto it = slot.find(indices[i]);
    if (it == slot.end()) {
      slot.emplace(indices[i], from.size());
      from.push_back(i);
      to.push_back(indices[i]);
    } else {
      from[it->second] = i;   // a later point overwrites the same field entry
    }
  }
  int prev = 0;
  int rc = check_hip(hipGetDevice(&prev), ctx);
  if (!rc) rc = check_hip(hipSetDevice(device), ctx);
  if (rc) return rc;
  nebo_point_map* m = new nebo_point_map();
  m->device = device;
  m->n = n;
  m->max_index = mx;
  m->m = (int64_t)from.size();
  m->gather = m->scatter_src = m->scatter_dst = nullptr;
  const size_t b1 = sizeof(int64_t) * (size_t)n, b2 = sizeof(int64_t) * from.size();
  if (n > 0) {
    rc = check_hip(hipMalloc((void**)&m->gather, b1 + 2 * b2), ctx);
    if (!rc) {
      m->scatter_src = m->gather + n;
      m->scatter_dst = m->scatter_src + m->m;
      rc = check_hip(hipMemcpy(m->gather, indices, b1, hipMemcpyHostToDevice), ctx);
    }
    if (!rc) rc = check_hip(hipMemcpy(m->scatter_src, from.data(), b2, hipMemcpyHostToDevice), ctx);
    if (!rc) rc = check_hip(hipMemcpy(m->scatter_dst, to.data(), b2, hipMemcpyHostToDevice), ctx);
  }
  (void)hipSetDevice(prev);
  if (rc) {
    if (m->gather) (void)hipFree(m->gather);
    delete m;
    return rc;
  }
  *out = m;
  return NEBO_OK;
}

extern "C" int nebo_point_map_destroy(nebo_point_map_t* map)
{
  clear_error();
  if (!map) return NEBO_OK;
  int rc = NEBO_OK;
  if (map->gather) {
    int prev = 0;
    rc = check_hip(hipGetDevice(&prev), "Nebo Point Field");
    if (!rc) rc = check_hip(hipSetDevice(map->device), "Nebo Point Field");
    if (!rc) rc = check_hip(hipFree(map->gather), "Nebo Point Field");
    (void)hipSetDevice(prev);
  }
  delete map;
  return rc;
}

extern "C" int nebo_point_to_field(const nebo_point_map_t* map, const nebo_field_t* dest, const double* src,
                                   void* stream)
{
  static const char* ctx = "Nebo Point Field";
  clear_error();
  int rc = check_map_field(map, dest, ctx);
  if (rc) return rc;
  if (map->m == 0) return NEBO_OK;
  if (!src) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((map->m + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, static_cast<double*>(dest->values), src, map->scatter_src,
                     map->scatter_dst, (long)map->m);
  return check_hip(hipGetLastError(), ctx);
}

extern "C" int nebo_field_to_point(const nebo_point_map_t* map, double* dest, const nebo_field_t* src,
                                   void* stream)
{
  static const char* ctx = "Nebo Point Field";
  clear_error();
  int rc = check_map_field(map, src, ctx);
  if (rc) return rc;
  if (map->n == 0) return NEBO_OK;
  if (!dest) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((map->n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, dest, static_cast<const double*>(src->values), map->gather,
                     (long)map->n);
  return check_hip(hipGetLastError(), ctx);
}
