// nebo_batch.hip -- many uneven patches per launch.
//
// Uintah hands SpatialOps one patch at a time; on a GPU that means one small
// launch per patch, each far too small to fill 256 CUs.  Here the patches live
// in a device table and the step is cut into work items (patch, x0, y0, z0) of a
// fixed tile shape, or (patch, q0, z0) runs of consecutive points of a patch's
// flattened x-y plane.  One persistent launch of ~4 workgroups per CU walks the item
// list with a fixed stride, so every patch of every size is covered by one
// launch, and a hipGraph holds several such steps.  The tile shape is chosen by
// timing candidates on the device (nebo_batch_autotune).
//
// The arithmetic per cell is the Laplacian + explicit Euler update of
// nebo_lapl_euler (out = in + s*((Lx + Ly) + Lz)), bit for bit.

#include "nebo_internal.h"
#include "nebo_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

namespace nebo {
namespace {

struct DevPatch {
  double* a;
  double* b;
  int nx, ny, nz;
  int X, XY;       // strides of the ghosted block
  int origin;      // flat index of interior cell (0,0,0)
};

struct Coef6 {
  double gx0, gx1, gy0, gy1, gz0, gz1, dx0, dx1, dy0, dy1, dz0, dz1;
};

__device__ __forceinline__ void patch_column(const double* __restrict__ in, double* __restrict__ out,
                                             int i, int XY, int X, int zn, const Coef6& c, double s)
{
  double pc = in[i];
  double gz_lo = in[i - XY] * c.gz0 + pc * c.gz1;
  for (int z = 0; z < zn; ++z, i += XY) {
    const double pz = in[i + XY];
    const double pxm = in[i - 1], pxp = in[i + 1], pym = in[i - X], pyp = in[i + X];
    const double gz_hi = pc * c.gz0 + pz * c.gz1;
    const double Lx = (pxm * c.gx0 + pc * c.gx1) * c.dx0 + (pc * c.gx0 + pxp * c.gx1) * c.dx1;
    const double Ly = (pym * c.gy0 + pc * c.gy1) * c.dy0 + (pc * c.gy0 + pyp * c.gy1) * c.dy1;
    const double Lz = gz_lo * c.dz0 + gz_hi * c.dz1;
    out[i] = pc + s * ((Lx + Ly) + Lz);
    pc = pz;
    gz_lo = gz_hi;
  }
}

// patch_column for a full chunk of TZ planes: every read of the chunk (TZ + 2
// planes of the column, 4 x/y neighbours per plane) is issued before the first
// cell is computed, so one memory latency covers TZ cells; the arithmetic is
// patch_column's, in the same order.
template <int TZ>
__device__ __forceinline__ void patch_chunk(const double* __restrict__ in, double* __restrict__ out, int i,
                                            int XY, int X, const Coef6& c, double s)
{
  double p[TZ + 2], xm[TZ], xp[TZ], ym[TZ], yp[TZ];
#pragma unroll
  for (int z = 0; z < TZ + 2; ++z) p[z] = in[i + (z - 1) * XY];
#pragma unroll
  for (int z = 0; z < TZ; ++z) {
    const int j = i + z * XY;
    xm[z] = in[j - 1];
    xp[z] = in[j + 1];
    ym[z] = in[j - X];
    yp[z] = in[j + X];
  }
  double gz_lo = p[0] * c.gz0 + p[1] * c.gz1;
#pragma unroll
  for (int z = 0; z < TZ; ++z) {
    const double pc = p[z + 1], pz = p[z + 2];
    const double gz_hi = pc * c.gz0 + pz * c.gz1;
    const double Lx = (xm[z] * c.gx0 + pc * c.gx1) * c.dx0 + (pc * c.gx0 + xp[z] * c.gx1) * c.dx1;
    const double Ly = (ym[z] * c.gy0 + pc * c.gy1) * c.dy0 + (pc * c.gy0 + yp[z] * c.gy1) * c.dy1;
    const double Lz = gz_lo * c.dz0 + gz_hi * c.dz1;
    out[i + z * XY] = pc + s * ((Lx + Ly) + Lz);
    gz_lo = gz_hi;
  }
}

// Flattened-plane items (tile y == 0): w = (patch, q0, -, z0) covers the tile-x
// consecutive points q0 .. q0 + tile_x - 1 of the patch's x-y plane (q = x + nx*y),
// so no lane idles on a patch narrower than the tile, and tz planes.  TZ > 0 is
// tz known at compile time: full chunks take patch_chunk, a chunk cut by the
// patch's end (uniform over the workgroup) the loop.
template <int TZ>
__global__ __launch_bounds__(1024) void batch_flat_kernel(const DevPatch* __restrict__ P,
                                                          const int4* __restrict__ items, int n_items,
                                                          int tz, Coef6 c, double s, int swap)
{
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 w = items[it];
    const DevPatch p = P[w.x];
    const int q = w.y + (int)threadIdx.x;
    if (q >= p.nx * p.ny) continue;
    const int y = q / p.nx, x = q - y * p.nx;
    const double* in = swap ? p.b : p.a;
    double* out = swap ? p.a : p.b;
    const int zn = min(tz, p.nz - w.w);
    const int i = p.origin + x + p.X * y + p.XY * w.w;
    if constexpr (TZ > 0) {
      if (zn == TZ) {
        patch_chunk<TZ>(in, out, i, p.XY, p.X, c, s);
        continue;
      }
    }
    patch_column(in, out, i, p.XY, p.X, zn, c, s);
  }
}

// Column items (tile x == 0, y = RY rows per lane, z = planes per item): one
// wave per item w = (patch, x0, y0, z0).  The wave covers a strip of up to 64
// columns of the patch from x0; a patch narrower than 64 columns puts
// G = min(64 / nx, 32 / RY) groups of RY rows side by side (lane = x + nx * g),
// so a 16-wide patch keeps every lane busy.  Each lane walks RY rows of its
// column through the item's planes, planes PF ahead held in registers: y
// neighbours are its own rows (plus one row above and below), z neighbours the
// neighbouring planes, x neighbours the adjacent lanes' values (DPP).  The
// values left and right of every group's rows (2 G RY <= 64 of them per plane)
// come from one load per plane, lane k loading value k, and reach the lanes at
// the groups' ends through ds_bpermute.  Loads and stores address a plane
// through a buffer resource (uniform base) and per-lane 32-bit offsets.  Four
// waves per workgroup, each its own item; the hardware hands workgroups to CUs as
// they free up, so uneven patches balance by themselves.  Workgroups are
// renumbered so that consecutive items run on one XCD.  Cells use
// patch_column's operations in its order.
__host__ __device__ constexpr int column_groups(int nx, int ry)
{
  return nx < kLanes ? (kLanes / nx < 32 / ry ? kLanes / nx : 32 / ry) : 1;
}

__device__ __forceinline__ double lane_gather(double v, int src)
{
  const int lo = __builtin_amdgcn_ds_bpermute(src * 4, __double2loint(v));
  const int hi = __builtin_amdgcn_ds_bpermute(src * 4, __double2hiint(v));
  return __hiloint2double(hi, lo);
}

template <int RY, int PF>
__global__ __launch_bounds__(256) void batch_column_kernel(const DevPatch* __restrict__ P,
                                                           const int4* __restrict__ items, int n_items, int zc,
                                                           Coef6 c, double s, int swap)
{
  static_assert(RY == 1 || RY == 2 || RY == 4, "rows per lane");
  // consecutive items (neighbouring rows, strips and chunks of a patch, which
  // share halo rows and planes) go to workgroups on the same XCD (one L2)
  const int it = __builtin_amdgcn_readfirstlane(xcd_tile((int)blockIdx.x, (int)gridDim.x) * 4 +
                                                (int)(threadIdx.x >> 6));
  if (it >= n_items) return;   // the whole wave: an item is one wave
  const int lane = (int)threadIdx.x & (kLanes - 1);
  const int4 w = items[it];
  const DevPatch p = P[w.x];
  const double* in = swap ? p.b : p.a;
  double* out = swap ? p.a : p.b;
  const int width = min(kLanes, p.nx - w.y);     // columns of this strip (== nx when G > 1)
  const int G = column_groups(p.nx, RY);
  const int g0 = lane / width;
  const bool active = g0 < G;
  const int g = min(g0, G - 1);
  const int xl = active ? lane - g * width : width - 1;
  const bool left = xl == 0, right = xl == width - 1;
  const int x = w.y + xl;
  const int yb = w.z + g * RY;                   // the lane's first row
  const int z0 = w.w, z1 = min(z0 + zc, p.nz);
  // byte offsets within a plane, from its point (-1, -1)
  auto off = [&](int xx, int yy) { return (uint32_t)((xx + 1) + p.X * (yy + 1)) * 8u; };
  uint32_t row[RY + 2];                          // rows yb-1 .. yb+RY, clamped to the ghost row
#pragma unroll
  for (int r = 0; r < RY + 2; ++r) row[r] = off(x, min(yb - 1 + r, p.ny));
  uint32_t edge;                                 // lane k < 2 G RY: value k = (group, row, side)
  {
    const int k = lane < 2 * G * RY ? lane : 0;
    const int ge = k / (2 * RY), je = (k >> 1) % RY;
    edge = off((k & 1) ? w.y + width : w.y - 1, min(w.z + ge * RY + je, p.ny));
  }
  const int src0 = 2 * g * RY;                   // this lane's group's first edge value

  auto plane = [&](const double* base, int z) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base) + (long)p.XY * (z + 1), (short)0, p.XY * 8,
                                             0x00020000);
  };
  struct Plane { double v[RY + 2], e; };
  auto load = [&](int z) {
    Plane q;
    const Rsrc rs = plane(in, z);
#pragma unroll
    for (int r = 0; r < RY + 2; ++r) q.v[r] = load_at(rs, row[r]);
    q.e = load_at(rs, edge);
    return q;
  };
  constexpr int NS = PF + 1;
  Plane S[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) S[q] = load(min(z0 + q, z1));
  double gz[RY];
  {
    const Rsrc rs = plane(in, z0 - 1);
#pragma unroll
    for (int j = 0; j < RY; ++j) gz[j] = load_at(rs, row[j + 1]) * c.gz0 + S[0].v[j + 1] * c.gz1;
  }
  auto step = [&](auto JJ, const int z) {
    constexpr int jj = decltype(JJ)::value;
    const Plane& C0 = S[jj];
    const Plane& C1 = S[(jj + 1) % NS];
    const Rsrc ro = plane(out, z);
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const double pc = C0.v[j + 1];
      const double lo = from_lower_lane(pc), hi = from_upper_lane(pc);
      const double el = lane_gather(C0.e, src0 + 2 * j), er = lane_gather(C0.e, src0 + 2 * j + 1);
      const double pxm = left ? el : lo, pxp = right ? er : hi;
      const double pym = C0.v[j], pyp = C0.v[j + 2];
      const double gz_hi = pc * c.gz0 + C1.v[j + 1] * c.gz1;
      const double Lx = (pxm * c.gx0 + pc * c.gx1) * c.dx0 + (pc * c.gx0 + pxp * c.gx1) * c.dx1;
      const double Ly = (pym * c.gy0 + pc * c.gy1) * c.dy0 + (pc * c.gy0 + pyp * c.gy1) * c.dy1;
      const double Lz = gz[j] * c.dz0 + gz_hi * c.dz1;
      gz[j] = gz_hi;
      if (active && yb + j < p.ny) store_nt_at(ro, row[j + 1], pc + s * ((Lx + Ly) + Lz));
    }
    // the slot's next plane is read by step z + NS - 1 at the earliest: past the
    // item's last plane (z1) no step reads it (z is uniform over the wave)
    if (z + NS <= z1) S[jj] = load(z + NS);
  };
  const int full = (z1 - z0) / NS, rem = (z1 - z0) - NS * full;
  int z = z0;
  for (int q = 0; q < full; ++q, z += NS)
    unroll_steps([&](auto J) { step(J, z + decltype(J)::value); }, std::make_integer_sequence<int, NS>{});
  unroll_steps([&](auto J) { if (rem > decltype(J)::value) step(J, z + decltype(J)::value); },
               std::make_integer_sequence<int, NS - 1>{});
}

// Work item w = (patch, x0, y0, z0).  Workgroup b takes items b, b + G, b + 2G, ...
__global__ __launch_bounds__(1024) void batch_kernel(const DevPatch* __restrict__ P,
                                                     const int4* __restrict__ items, int n_items,
                                                     int tx, int ty, int tz, Coef6 c, double s,
                                                     int swap)
{
  const int lx = threadIdx.x % tx;
  const int ly = threadIdx.x / tx;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int4 w = items[it];
    const DevPatch p = P[w.x];
    const int x = w.y + lx, y = w.z + ly;
    if (x >= p.nx || y >= p.ny) continue;
    const double* in = swap ? p.b : p.a;
    double* out = swap ? p.a : p.b;
    const int zn = min(tz, p.nz - w.w);
    patch_column(in, out, p.origin + x + p.X * y + p.XY * w.w, p.XY, p.X, zn, c, s);
  }
}

Coef6 coef6(const nebo_axis_coefs_t& c)
{
  return Coef6{c.grad[0][0], c.grad[0][1], c.grad[1][0], c.grad[1][1], c.grad[2][0], c.grad[2][1],
               c.div[0][0],  c.div[0][1],  c.div[1][0],  c.div[1][1],  c.div[2][0],  c.div[2][1]};
}

}  // namespace
}  // namespace nebo

using namespace nebo;

struct nebo_batch {
  std::vector<nebo_patch_t> patches;
  nebo_axis_coefs_t coefs;
  double s = 0.0;
  int tile[3] = {64, 4, 16};
  int device = 0;
  int cus = 256;
  DevPatch* d_patches = nullptr;
  int4* d_items = nullptr;
  int n_items = 0;
  std::vector<int> patch_first;   // first item of each patch (items are grouped by patch)
  int blocks = 0;
  long max_plane = 0;   // largest (nx+2)*(ny+2) of the patches
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int build_items(nebo_batch* b)
{
  static const char* ctx = "Nebo Batch";
  std::vector<int4> items;
  b->patch_first.assign(b->patches.size() + 1, 0);
  const int tx = b->tile[0], ty = b->tile[1], tz = b->tile[2];
  for (size_t k = 0; k < b->patches.size(); ++k) {
    b->patch_first[k] = (int)items.size();
    const nebo_patch_t& p = b->patches[k];
    for (int z = 0; z < p.n[2]; z += tz) {
      if (tx == 0) {   // column schedule: strips of 64 columns, row groups of 64/nx x ty rows
        const int rows = column_groups(p.n[0], ty) * ty;
        for (int y = 0; y < p.n[1]; y += rows)
          for (int x = 0; x < p.n[0]; x += kLanes) items.push_back(make_int4((int)k, x, y, z));
      } else if (ty == 0) {   // flattened plane
        for (int q = 0; q < p.n[0] * p.n[1]; q += tx) items.push_back(make_int4((int)k, q, 0, z));
      } else {
        for (int y = 0; y < p.n[1]; y += ty)
          for (int x = 0; x < p.n[0]; x += tx) items.push_back(make_int4((int)k, x, y, z));
      }
    }
  }
  b->patch_first[b->patches.size()] = (int)items.size();
  if (b->d_items) {
    // launches still in flight may read the old table
    int rc0 = check_hip(hipDeviceSynchronize(), ctx);
    if (rc0) return rc0;
    (void)hipFree(b->d_items);
  }
  b->d_items = nullptr;
  b->n_items = (int)items.size();
  int rc = check_hip(hipMalloc((void**)&b->d_items, sizeof(int4) * std::max<size_t>(1, items.size())), ctx);
  if (rc) return rc;
  rc = check_hip(hipMemcpy(b->d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice), ctx);
  if (rc) return rc;
  if (tx == 0) {   // one wave per item, four per workgroup
    b->blocks = std::max(1, (b->n_items + 3) / 4);
    return NEBO_OK;
  }
  // persistent grid: a few workgroups per CU, never more than there are items
  const int threads = ty == 0 ? tx : tx * ty;
  const int per_cu = threads <= 256 ? 8 : threads <= 512 ? 4 : 2;
  b->blocks = std::max(1, std::min(b->n_items, b->cus * per_cu));
  return NEBO_OK;
}

void drop_graph(nebo_batch* b)
{
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int check_tile(const int32_t* t)
{
  if (t[0] == 0) {
    if ((t[1] != 1 && t[1] != 2 && t[1] != 4) || t[2] < 1)
      return set_error(NEBO_ERR_INVALID, "Nebo Batch",
                       "a column tile (x = 0) takes 1, 2 or 4 rows per lane (y) and z >= 1");
    return NEBO_OK;
  }
  if (t[1] == 0) {
    if (t[0] < 64 || t[0] > 1024 || t[0] % 64 || t[2] < 1)
      return set_error(NEBO_ERR_INVALID, "Nebo Batch",
                       "a flattened tile (y = 0) takes a multiple of 64 threads up to 1024 and z >= 1");
    return NEBO_OK;
  }
  if (t[0] < 1 || t[1] < 1 || t[2] < 1 || t[0] * t[1] > 1024)
    return set_error(NEBO_ERR_INVALID, "Nebo Batch", "tile must be positive with x*y <= 1024 threads");
  return NEBO_OK;
}

// The column kernel for ry rows per lane and items of zc planes.  Items of up to
// 4 or 8 planes load every plane they read as they start (planes ahead >= item
// length), so a wave has all its loads in flight at once; longer items keep 2-3
// planes ahead.  512 patches of 16-64^3 (hipGraph of 10 steps, one box): items of
// 4 planes with 2 rows per lane 0.115 ms per step, of 8 planes 0.117, against 0.131
// for 16-plane items 3 planes ahead; 4-plane items took 0.151 ms 3 planes ahead and
// 0.121-0.155 ms 5-9 ahead (whose first loads repeat the item's last plane).
using ColumnKernel = void (*)(const DevPatch*, const int4*, int, int, Coef6, double, int);
ColumnKernel column_kernel(int ry, int zc)
{
  if (zc <= 4) return ry == 1 ? batch_column_kernel<1, 4> : ry == 2 ? batch_column_kernel<2, 4> : batch_column_kernel<4, 4>;
  if (zc <= 8) return ry == 1 ? batch_column_kernel<1, 8> : ry == 2 ? batch_column_kernel<2, 8> : batch_column_kernel<4, 8>;
  return ry == 1 ? batch_column_kernel<1, 3> : ry == 2 ? batch_column_kernel<2, 3> : batch_column_kernel<4, 2>;
}

int launch_items(nebo_batch* b, int first, int count, int grid, int swap, hipStream_t s)
{
  if (count <= 0) return NEBO_OK;
  if (b->tile[0] == 0) {
    const ColumnKernel fn = column_kernel(b->tile[1], b->tile[2]);
    hipLaunchKernelGGL(fn, dim3((count + 3) / 4), dim3(256), 0, s, b->d_patches, b->d_items + first, count,
                       b->tile[2], coef6(b->coefs), b->s, swap);
  } else if (b->tile[1] == 0) {
    const int tz = b->tile[2];
    auto fn = tz == 4 ? batch_flat_kernel<4> : tz == 8 ? batch_flat_kernel<8> : tz == 16 ? batch_flat_kernel<16>
                                                                                         : batch_flat_kernel<0>;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(b->tile[0]), 0, s, b->d_patches, b->d_items + first, count, tz,
                       coef6(b->coefs), b->s, swap);
  } else {
    hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(b->tile[0] * b->tile[1]), 0, s, b->d_patches,
                       b->d_items + first, count, b->tile[0], b->tile[1], b->tile[2], coef6(b->coefs), b->s,
                       swap);
  }
  return check_hip(hipGetLastError(), "Nebo Batch");
}

}  // namespace

extern "C" int nebo_batch_create(const nebo_patch_t* patches, int32_t n, const nebo_axis_coefs_t* c,
                                 double s, const int32_t* tile, nebo_batch_t** out)
{
  static const char* ctx = "Nebo Batch";
  clear_error();
  if (!patches || n <= 0 || !c || !out) return set_error(NEBO_ERR_INVALID, ctx, "null argument or no patches");
  *out = nullptr;
  if (tile) {
    int rc = check_tile(tile);
    if (rc) return rc;
  }
  for (int i = 0; i < n; ++i) {
    const nebo_patch_t& p = patches[i];
    if (!p.a || !p.b || p.a == p.b || p.n[0] < 1 || p.n[1] < 1 || p.n[2] < 1)
      return set_error(NEBO_ERR_INVALID, ctx, "patch needs two distinct blocks and positive sizes");
    if ((long)(p.n[0] + 2) * (p.n[1] + 2) * (p.n[2] + 2) >= (1L << 31))
      return set_error(NEBO_ERR_TOO_LARGE, ctx, "patch has 2^31 or more points");
  }
  long max_plane = 0;
  for (int i = 0; i < n; ++i)
    max_plane = std::max(max_plane, (long)(patches[i].n[0] + 2) * (patches[i].n[1] + 2));
  // the column schedule addresses a plane with 32-bit byte offsets
  if (tile && tile[0] == 0 && max_plane * 8 >= (1L << 31))
    return set_error(NEBO_ERR_TOO_LARGE, ctx, "a column tile needs every patch plane below 2^31 bytes");
  nebo_batch* b = new nebo_batch;
  b->max_plane = max_plane;
  b->patches.assign(patches, patches + n);
  b->coefs = *c;
  b->s = s;
  if (tile) std::memcpy(b->tile, tile, sizeof b->tile);
  int rc = check_hip(hipGetDevice(&b->device), ctx);
  hipDeviceProp_t prop;
  if (!rc) rc = check_hip(hipGetDeviceProperties(&prop, b->device), ctx);
  if (!rc) b->cus = prop.multiProcessorCount;
  std::vector<DevPatch> dp(n);
  for (int i = 0; i < n; ++i) {
    const nebo_patch_t& p = patches[i];
    DevPatch& d = dp[i];
    d.a = p.a;
    d.b = p.b;
    d.nx = p.n[0];
    d.ny = p.n[1];
    d.nz = p.n[2];
    d.X = p.n[0] + 2;
    d.XY = d.X * (p.n[1] + 2);
    d.origin = 1 + d.X + d.XY;
  }
  if (!rc) rc = check_hip(hipMalloc((void**)&b->d_patches, sizeof(DevPatch) * n), ctx);
  if (!rc) rc = check_hip(hipMemcpy(b->d_patches, dp.data(), sizeof(DevPatch) * n, hipMemcpyHostToDevice), ctx);
  if (!rc) rc = build_items(b);
  if (rc) {
    nebo_batch_destroy(b);
    return rc;
  }
  *out = b;
  return NEBO_OK;
}

extern "C" int nebo_batch_destroy(nebo_batch_t* b)
{
  clear_error();
  if (!b) return NEBO_OK;
  drop_graph(b);
  if (b->d_patches) (void)hipFree(b->d_patches);
  if (b->d_items) (void)hipFree(b->d_items);
  delete b;
  return NEBO_OK;
}

extern "C" int nebo_batch_step(nebo_batch_t* b, int swap, void* stream)
{
  clear_error();
  if (!b) return set_error(NEBO_ERR_INVALID, "Nebo Batch", "null batch");
  return launch_items(b, 0, b->n_items, b->blocks, swap ? 1 : 0, (hipStream_t)stream);
}

extern "C" int nebo_batch_step_per_patch(nebo_batch_t* b, int swap, void* stream)
{
  clear_error();
  if (!b) return set_error(NEBO_ERR_INVALID, "Nebo Batch", "null batch");
  for (size_t k = 0; k < b->patches.size(); ++k) {
    const int first = b->patch_first[k], count = b->patch_first[k + 1] - first;
    int rc = launch_items(b, first, count, count, swap ? 1 : 0, (hipStream_t)stream);
    if (rc) return rc;
  }
  return NEBO_OK;
}

extern "C" int nebo_batch_graph_build(nebo_batch_t* b, int32_t steps, void* stream)
{
  static const char* ctx = "Nebo Batch";
  clear_error();
  if (!b || steps < 1 || !stream) return set_error(NEBO_ERR_INVALID, ctx, "null batch/stream or no steps");
  drop_graph(b);
  hipStream_t s = (hipStream_t)stream;
  int rc = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), ctx);
  if (rc) return rc;
  for (int i = 0; i < steps && !rc; ++i) rc = launch_items(b, 0, b->n_items, b->blocks, i & 1, s);
  hipGraph_t g = nullptr;
  const int rc2 = check_hip(hipStreamEndCapture(s, &g), ctx);
  if (rc || rc2) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : rc2;
  }
  b->graph = g;
  return check_hip(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), ctx);
}

extern "C" int nebo_batch_graph_launch(nebo_batch_t* b, void* stream)
{
  clear_error();
  if (!b || !b->exec) return set_error(NEBO_ERR_INVALID, "Nebo Batch", "no graph built");
  return check_hip(hipGraphLaunch(b->exec, (hipStream_t)stream), "Nebo Batch");
}

extern "C" int nebo_batch_autotune(nebo_batch_t* b, int32_t* tile_out, void* stream)
{
  static const char* ctx = "Nebo Batch";
  clear_error();
  if (!b) return set_error(NEBO_ERR_INVALID, ctx, "null batch");
  drop_graph(b);
  hipStream_t s = (hipStream_t)stream;
  static const int cand[][3] = {
      {64, 4, 16}, {64, 4, 32}, {64, 4, 8}, {64, 2, 16}, {64, 8, 16}, {64, 8, 8}, {64, 16, 8},
      {32, 8, 16}, {32, 8, 8}, {32, 4, 16}, {32, 16, 8}, {16, 16, 16}, {16, 16, 8}, {16, 8, 16},
      {64, 1, 64}, {128, 2, 16}, {128, 4, 8}, {256, 1, 16},
      // flattened plane (y = 0)
      {64, 0, 16}, {128, 0, 8}, {128, 0, 16}, {256, 0, 4}, {256, 0, 8}, {256, 0, 16}, {512, 0, 4},
      {512, 0, 8}, {1024, 0, 4}, {1024, 0, 8},
      // column schedule (x = 0): rows per lane, planes per item
      {0, 1, 16}, {0, 1, 32}, {0, 1, 64}, {0, 2, 16}, {0, 2, 32}, {0, 2, 64}, {0, 4, 16}, {0, 4, 32},
      {0, 4, 64}, {0, 1, 8}, {0, 2, 8}, {0, 4, 8}, {0, 1, 4}, {0, 2, 4}, {0, 4, 4}};
  hipEvent_t e0, e1;
  int rc = check_hip(hipEventCreate(&e0), ctx);
  if (!rc) rc = check_hip(hipEventCreate(&e1), ctx);
  if (rc) return rc;
  float best = 1e30f;
  int best_tile[3] = {b->tile[0], b->tile[1], b->tile[2]};
  for (const auto& t : cand) {
    if (t[0] == 0 && b->max_plane * 8 >= (1L << 31)) continue;   // planes too large for 32-bit offsets
    std::memcpy(b->tile, t, sizeof b->tile);
    if ((rc = build_items(b))) break;
    // warm once, then time four steps in the order a time loop runs them
    // (a -> b, b -> a, ...: each step reads what the one before wrote)
    if ((rc = launch_items(b, 0, b->n_items, b->blocks, 1, s))) break;
    if ((rc = check_hip(hipEventRecord(e0, s), ctx))) break;
    for (int r = 0; r < 4 && !rc; ++r) rc = launch_items(b, 0, b->n_items, b->blocks, r & 1, s);
    if (rc || (rc = check_hip(hipEventRecord(e1, s), ctx))) break;
    if ((rc = check_hip(hipEventSynchronize(e1), ctx))) break;
    float ms = 0;
    if ((rc = check_hip(hipEventElapsedTime(&ms, e0, e1), ctx))) break;
    if (ms < best) {
      best = ms;
      std::memcpy(best_tile, t, sizeof best_tile);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::memcpy(b->tile, best_tile, sizeof b->tile);
  const int rc2 = build_items(b);
  if (rc) return rc;
  if (rc2) return rc2;
  if (tile_out) std::memcpy(tile_out, b->tile, sizeof b->tile);
  return NEBO_OK;
}

extern "C" int nebo_batch_info(nebo_batch_t* b, int32_t* tile_out, int32_t* n_items, int32_t* n_blocks)
{
  clear_error();
  if (!b) return set_error(NEBO_ERR_INVALID, "Nebo Batch", "null batch");
  if (tile_out) std::memcpy(tile_out, b->tile, sizeof b->tile);
  if (n_items) *n_items = b->n_items;
  if (n_blocks) *n_blocks = b->blocks;
  return NEBO_OK;
}
