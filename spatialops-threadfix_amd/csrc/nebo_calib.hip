// nebo_calib.hip -- streaming kernels that calibrate the roofline of the stencil
// kernels (no counterpart in the reference).
//
// r = a + s*b over n doubles: two read streams and one write stream, the
// traffic mix of div(k grad phi) (read phi and k, write r) without its halos or
// arithmetic.  Its rate on the same blocks is the streaming ceiling the stencil
// kernel's HBM rate is measured against.  Variants differ in access width and
// cache policy:
//   0: 8-byte accesses, nontemporal stores
//   1: 16-byte accesses, nontemporal stores
//   2: 16-byte accesses, default (write-back) stores
//   3: 32-byte accesses (two 16-byte accesses per lane and array), nontemporal stores
//   4: 16-byte accesses at an 8-byte offset (the arrays start 8 bytes past a
//      16-byte boundary, as the interior rows of a field with one ghost layer do),
//      nontemporal stores
// Each thread issues all its loads before its first store; blocks take
// consecutive chunks, renumbered so that consecutive chunks run on one XCD.

#include "nebo_internal.h"
#include "nebo_device.h"

#include <hip/hip_runtime.h>

namespace nebo {
namespace {

constexpr int kCalibBlock = 256;

using D2 = __attribute__((__vector_size__(2 * sizeof(double)))) double;

template <int VEC, bool NT>
__global__ __launch_bounds__(kCalibBlock) void triad_kernel(double* __restrict__ r, const double* __restrict__ a,
                                                            const double* __restrict__ b, double s, long n)
{
  const long chunk = (long)kCalibBlock * VEC;
  const long base = (long)xcd_tile(blockIdx.x, gridDim.x) * chunk;
  if constexpr (VEC == 1) {
    const long i = base + threadIdx.x;
    if (i >= n) return;
    const double v = a[i] + s * b[i];
    if constexpr (NT) __builtin_nontemporal_store(v, &r[i]);
    else r[i] = v;
  } else {
    // VEC / 2 pairs per thread, pair p of thread t at base + 2*(p*kCalibBlock + t)
    constexpr int P = VEC / 2;
    D2 va[P], vb[P];
    long idx[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      idx[p] = base + 2L * ((long)p * kCalibBlock + threadIdx.x);
      if (idx[p] + 1 < n) {
        va[p] = *reinterpret_cast<const D2*>(a + idx[p]);
        vb[p] = *reinterpret_cast<const D2*>(b + idx[p]);
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (idx[p] + 1 < n) {
        D2 v;
        v[0] = va[p][0] + s * vb[p][0];
        v[1] = va[p][1] + s * vb[p][1];
        if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<D2*>(r + idx[p]));
        else *reinterpret_cast<D2*>(r + idx[p]) = v;
      } else if (idx[p] < n) {
        r[idx[p]] = a[idx[p]] + s * b[idx[p]];
      }
    }
  }
}

}  // namespace
}  // namespace nebo

using namespace nebo;

extern "C" int nebo_stream_triad(double* r, const double* a, const double* b, double s, int64_t n, int32_t variant,
                                 void* stream)
{
  static const char* ctx = "Nebo Calibration";
  clear_error();
  if (!r || !a || !b || n < 0) return set_error(NEBO_ERR_INVALID, ctx, "null argument or negative size");
  if (variant < 0 || variant > 4) return set_error(NEBO_ERR_INVALID, ctx, "triad variant must be 0..4");
  if (n == 0) return NEBO_OK;
  const int vec = variant == 0 ? 1 : variant == 3 ? 4 : 2;
  const uintptr_t mis = (reinterpret_cast<uintptr_t>(r) | reinterpret_cast<uintptr_t>(a) |
                         reinterpret_cast<uintptr_t>(b)) & 15;
  if (variant >= 1 && variant <= 3 && mis)
    return set_error(NEBO_ERR_INVALID, ctx, "16-byte variants 1-3 need 16-byte aligned arrays");
  if (variant == 4 && ((reinterpret_cast<uintptr_t>(r) & 15) != 8 || (reinterpret_cast<uintptr_t>(a) & 15) != 8 ||
                       (reinterpret_cast<uintptr_t>(b) & 15) != 8))
    return set_error(NEBO_ERR_INVALID, ctx, "variant 4 needs arrays 8 bytes past a 16-byte boundary");
  const long blocks = (n + (long)kCalibBlock * vec - 1) / ((long)kCalibBlock * vec);
  if (blocks >= (1L << 31)) return set_error(NEBO_ERR_TOO_LARGE, ctx, "more than 2^31 blocks");
  hipStream_t st = (hipStream_t)stream;
  switch (variant) {
  case 0: hipLaunchKernelGGL((triad_kernel<1, true>), dim3(blocks), dim3(kCalibBlock), 0, st, r, a, b, s, n); break;
  case 1: hipLaunchKernelGGL((triad_kernel<2, true>), dim3(blocks), dim3(kCalibBlock), 0, st, r, a, b, s, n); break;
  case 2: hipLaunchKernelGGL((triad_kernel<2, false>), dim3(blocks), dim3(kCalibBlock), 0, st, r, a, b, s, n); break;
  case 3: hipLaunchKernelGGL((triad_kernel<4, true>), dim3(blocks), dim3(kCalibBlock), 0, st, r, a, b, s, n); break;
  default: hipLaunchKernelGGL((triad_kernel<2, true>), dim3(blocks), dim3(kCalibBlock), 0, st, r, a, b, s, n); break;
  }
  return check_hip(hipGetLastError(), ctx);
}
