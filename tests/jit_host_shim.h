// Host stand-ins for the device builtins that the engine's generated column-schedule
// kernels use, so that a generated kernel's address arithmetic runs on the CPU,
// one block and one lane at a time, with every memory access checked against
// the allocations a test registers (see tests/test_fused_addresses.py).
//
// Only addresses are audited: lane exchanges and lane broadcasts return the
// lane's own value, so the computed values are meaningless.  A buffer load whose
// offset lies past the resource's range reads nothing (the hardware returns 0);
// one inside the range must address registered memory, since the hardware reads
// memory there.  Stores are checked and counted, not performed.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(n)

struct ShimDim {
  unsigned x = 0, y = 0, z = 0;
};
static ShimDim blockIdx, threadIdx;

struct ShimRange {
  const char* lo;
  const char* hi;
};
static ShimRange g_ranges[64];
static int g_nranges = 0;
static long g_bad = 0, g_loads = 0, g_stores = 0;
static const char* g_first_bad = nullptr;

static inline bool shim_inside(const void* p, size_t bytes)
{
  const char* c = static_cast<const char*>(p);
  for (int i = 0; i < g_nranges; ++i)
    if (c >= g_ranges[i].lo && c + bytes <= g_ranges[i].hi) return true;
  return false;
}

static inline void shim_bad(const void* p)
{
  ++g_bad;
  if (!g_first_bad) g_first_bad = static_cast<const char*>(p);
}

struct ShimRsrc {
  const char* base;
  long n;
};
#define __amdgpu_buffer_rsrc_t ShimRsrc

static inline ShimRsrc shim_make_rsrc(void* p, short, int n, int) { return ShimRsrc{static_cast<const char*>(p), n}; }
#define __builtin_amdgcn_make_buffer_rsrc shim_make_rsrc

typedef unsigned ShimU2 __attribute__((vector_size(8)));
static inline ShimU2 shim_load_b64(ShimRsrc r, unsigned off, int, int)
{
  ShimU2 v = {0u, 0u};
  ++g_loads;
  if ((long)off + 8 <= r.n) {   // inside the range: the hardware reads this memory
    const char* a = r.base + off;
    if (!shim_inside(a, 8)) shim_bad(a);
    else std::memcpy(&v, a, 8);
  }
  return v;
}
#define __builtin_amdgcn_raw_buffer_load_b64 shim_load_b64

// A 16-byte load (two adjacent doubles): each 8-byte half inside the range must
// address registered memory (a half past the range reads nothing).
typedef unsigned ShimU4 __attribute__((vector_size(16)));
static inline ShimU4 shim_load_b128(ShimRsrc r, unsigned off, int, int)
{
  ShimU4 v = {0u, 0u, 0u, 0u};
  ++g_loads;
  for (int q = 0; q < 2; ++q) {
    const long o = (long)off + 8 * q;
    if (o + 8 > r.n) continue;
    const char* a = r.base + o;
    if (!shim_inside(a, 8)) {
      shim_bad(a);
      continue;
    }
    unsigned w[2];
    std::memcpy(w, a, 8);
    v[2 * q] = w[0];
    v[2 * q + 1] = w[1];
  }
  return v;
}
#define __builtin_amdgcn_raw_buffer_load_b128 shim_load_b128
// the generated kernels' 16-byte value type (two doubles, 8-byte aligned)
typedef double ShimD2 __attribute__((vector_size(16), aligned(8)));
#define NEBO_JIT_D2 ShimD2

static inline int shim_readlane(int v, int) { return v; }
#define __builtin_amdgcn_readlane shim_readlane
static inline double shim_dpp(double v, int, int, int, bool) { return v; }
#define __builtin_amdgcn_mov_dpp shim_dpp

static inline int __double2loint(double d)
{
  int r[2];
  std::memcpy(r, &d, 8);
  return r[0];
}
static inline int __double2hiint(double d)
{
  int r[2];
  std::memcpy(r, &d, 8);
  return r[1];
}
static inline double __hiloint2double(int hi, int lo)
{
  int r[2] = {lo, hi};
  double d;
  std::memcpy(&d, r, 8);
  return d;
}

// Stores go to a hook the audit driver defines (it checks and counts them).
static void shim_store_hit(const void* p, size_t bytes);
#define __builtin_nontemporal_store(v, p) (++g_stores, shim_store_hit((p), sizeof(*(p))))

// device math the generated code may call (values are not audited)
static inline double erfinv(double x) { return x; }
static inline double erfcinv(double x) { return x; }
using std::cos;
using std::erf;
using std::erfc;
using std::exp;
using std::log;
using std::log10;
using std::pow;
using std::sin;
using std::tan;
using std::tanh;
