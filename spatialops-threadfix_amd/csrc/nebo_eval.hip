// nebo_eval.hip -- the generic fused evaluator and the assignment / reduction
// entry points.
//
// An expression tree is expanded on the host into a post-order stack program in
// which every stencil point has been folded into the offsets of the loads
// (nebo_lower.cpp).  The device walks that program once per cell.  The stack
// depth before every instruction is known on the host and is part of the
// opcode, so each (opcode, depth) pair is a separate case with fixed stack
// registers: there is no dynamically indexed register file and no scratch.
// The program lives in the kernel-argument segment, so instruction fetches are
// uniform scalar loads and every branch is uniform across the wavefront.
//
// Each thread evaluates V consecutive z cells of one (x, y) column; the
// (x, y) plane is flattened so that a wavefront covers 64 consecutive x (or
// the rows of a narrow box) and loads stay coalesced.

#include "nebo_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <unordered_map>

namespace nebo {
namespace {

constexpr int kBlock = 256;
constexpr int kVmV = 2;           // z cells per thread in the box modes
constexpr int kRedMaxBlocks = 2048;

// A reduction's working memory: 2 * kRedMaxBlocks partial doubles and a result
// slot, kRedBufBytes in all.
struct RedBuf {
  double* partial;
  double* result;
};
constexpr size_t kRedBufBytes = sizeof(double) * (2 * kRedMaxBlocks + 1);
inline RedBuf red_buf_at(char* p)
{
  double* d = reinterpret_cast<double*>(p);
  return RedBuf{d, d + 2 * kRedMaxBlocks};
}

}  // namespace
}  // namespace nebo

// Working memory of nebo_reduce_async_noalloc: the block partials and result
// slot of one reduction (nebo::kRedBufBytes = 32 KB), one copy per GPU, held
// until the process ends.  Host code never passes this name to a kernel: on the
// host it is only a handle for hipGetSymbolAddress, which gives each device's
// address (nebo::scratch_on_device).
static __device__ __align__(256) char scratch[nebo::kRedBufBytes];

namespace nebo {
namespace {

__device__ __noinline__ double vm_math(int op, double x)
{
  switch (op) {
  case VM_EXP: return exp(x);
  case VM_LOG: return log(x);
  case VM_LOG10: return log10(x);
  case VM_SIN: return sin(x);
  case VM_COS: return cos(x);
  case VM_TAN: return tan(x);
  case VM_TANH: return tanh(x);
  case VM_ERF: return erf(x);
  case VM_ERFC: return erfc(x);
  case VM_INV_ERF: return erfinv(x);
  case VM_INV_ERFC: return erfcinv(x);
  default: return x;
  }
}

__device__ __noinline__ double vm_pow(double a, double b) { return pow(a, b); }

template <int NCLS, int V>
struct CellIdx {
  int i[NCLS][V];   // flat index of each cell in each glob-shape class
  int x, y;
  int z[V];
};

template <int NCLS, int V>
__device__ __forceinline__ int pick(const CellIdx<NCLS, V>& c, int cls, int v)
{
  if constexpr (NCLS == 1) {
    return c.i[0][v];
  } else {
    int r = c.i[0][v];
#pragma unroll
    for (int k = 1; k < NCLS; ++k) r = (cls == k) ? c.i[k][v] : r;
    return r;
  }
}

template <int OP>
constexpr int vm_pops()
{
  if constexpr (OP == VM_LOAD || OP == VM_LOADG || OP == VM_CONST || OP == VM_MASK) return 0;
  else if constexpr (OP >= VM_ADD && OP <= VM_OR) return 2;
  else if constexpr (OP == VM_SELECT) return 3;
  else return 1;
}

template <int OP, int D, int NCLS, int V, bool MATH>
__device__ __forceinline__ void vm_exec(double (&S)[kVmMaxDepth][V], const VmProgram& P,
                                        int imm, const CellIdx<NCLS, V>& c)
{
  constexpr int pops = vm_pops<OP>();
  constexpr bool push = (OP == VM_LOAD || OP == VM_LOADG || OP == VM_CONST || OP == VM_MASK);
  constexpr bool math = (OP == VM_POW || OP == VM_POWC || (OP >= VM_EXP && OP <= VM_INV_ERFC));
  if constexpr (D < pops || (push && D >= kVmMaxDepth) || (math && !MATH)) {
    return;  // never emitted by the host for this instantiation
  } else if constexpr (OP == VM_LOAD) {
    const double* b = P.loads[imm].base;
    const int cls = P.loads[imm].cls;
#pragma unroll
    for (int v = 0; v < V; ++v) S[D][v] = b[pick(c, cls, v)];
  } else if constexpr (OP == VM_LOADG) {
    // an edgeless stencil may read past its operand's block at cells that a
    // cond does not select; those reads yield NaN instead of touching memory
    const VmLoad& L = P.loads[imm];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int i = pick(c, L.cls, v);
      S[D][v] = (i >= L.glo && i < L.ghi) ? L.base[i] : __builtin_nan("");
    }
  } else if constexpr (OP == VM_CONST) {
    const double k = P.consts[imm];
#pragma unroll
    for (int v = 0; v < V; ++v) S[D][v] = k;
  } else if constexpr (OP == VM_MASK) {
    const VmMask& m = P.masks[imm];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int pos = (m.o[0] + c.x) + m.X * ((m.o[1] + c.y) + m.Y * (m.o[2] + c.z[v]));
      S[D][v] = ((m.bits[pos >> 5] >> (pos & 31)) & 1u) ? 1.0 : 0.0;
    }
  } else if constexpr (OP == VM_MULC) {
    const double k = P.consts[imm];
#pragma unroll
    for (int v = 0; v < V; ++v) S[D - 1][v] = S[D - 1][v] * k;
  } else if constexpr (pops == 2) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const double a = S[D - 2][v], b = S[D - 1][v];
      double r;
      if constexpr (OP == VM_ADD) r = a + b;
      else if constexpr (OP == VM_SUB) r = a - b;
      else if constexpr (OP == VM_MUL) r = a * b;
      else if constexpr (OP == VM_DIV) r = a / b;
      else if constexpr (OP == VM_MAX) r = (a > b) ? a : b;
      else if constexpr (OP == VM_MIN) r = (a < b) ? a : b;
      else if constexpr (OP == VM_POW) r = vm_pow(a, b);
      else if constexpr (OP == VM_EQ) r = (a == b) ? 1.0 : 0.0;
      else if constexpr (OP == VM_NE) r = (a != b) ? 1.0 : 0.0;
      else if constexpr (OP == VM_LT) r = (a < b) ? 1.0 : 0.0;
      else if constexpr (OP == VM_LE) r = (a <= b) ? 1.0 : 0.0;
      else if constexpr (OP == VM_GT) r = (a > b) ? 1.0 : 0.0;
      else if constexpr (OP == VM_GE) r = (a >= b) ? 1.0 : 0.0;
      else if constexpr (OP == VM_AND) r = (a != 0.0 && b != 0.0) ? 1.0 : 0.0;
      else r = (a != 0.0 || b != 0.0) ? 1.0 : 0.0;  // VM_OR
      S[D - 2][v] = r;
    }
  } else if constexpr (OP == VM_SELECT) {
#pragma unroll
    for (int v = 0; v < V; ++v)
      S[D - 3][v] = (S[D - 3][v] != 0.0) ? S[D - 2][v] : S[D - 1][v];
  } else {
    // unary
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const double a = S[D - 1][v];
      double r;
      if constexpr (OP == VM_NEG) r = -a;
      else if constexpr (OP == VM_ABS) r = fabs(a);
      else if constexpr (OP == VM_SQRT) r = __builtin_sqrt(a);
      else if constexpr (OP == VM_SQUARE) r = a * a;
      else if constexpr (OP == VM_CUBE) r = a * a * a;
      else if constexpr (OP == VM_POWC) r = vm_pow(a, P.consts[imm]);
      else if constexpr (OP == VM_NOT) r = (a != 0.0) ? 0.0 : 1.0;
      else r = vm_math(OP, a);
      S[D - 1][v] = r;
    }
  }
}

#define VM_C(OP, D) \
  case ((OP) << 4) | (D): vm_exec<OP, D, NCLS, V, MATH>(S, P, imm, c); break;
#define VM_CD(OP)                                                                      \
  VM_C(OP, 0) VM_C(OP, 1) VM_C(OP, 2) VM_C(OP, 3) VM_C(OP, 4) VM_C(OP, 5) VM_C(OP, 6) \
  VM_C(OP, 7) VM_C(OP, 8) VM_C(OP, 9) VM_C(OP, 10)
#define VM_ALL_CASES                                                                    \
  VM_CD(VM_LOAD) VM_CD(VM_CONST) VM_CD(VM_MASK) VM_CD(VM_MULC) VM_CD(VM_ADD)           \
  VM_CD(VM_SUB) VM_CD(VM_MUL) VM_CD(VM_DIV) VM_CD(VM_MAX) VM_CD(VM_MIN) VM_CD(VM_POW)  \
  VM_CD(VM_EQ) VM_CD(VM_NE) VM_CD(VM_LT) VM_CD(VM_LE) VM_CD(VM_GT) VM_CD(VM_GE)        \
  VM_CD(VM_AND) VM_CD(VM_OR) VM_CD(VM_NEG) VM_CD(VM_ABS) VM_CD(VM_SQRT)                \
  VM_CD(VM_SQUARE) VM_CD(VM_CUBE) VM_CD(VM_POWC) VM_CD(VM_EXP) VM_CD(VM_LOG)           \
  VM_CD(VM_LOG10) VM_CD(VM_SIN) VM_CD(VM_COS) VM_CD(VM_TAN) VM_CD(VM_TANH)             \
  VM_CD(VM_ERF) VM_CD(VM_ERFC) VM_CD(VM_INV_ERF) VM_CD(VM_INV_ERFC) VM_CD(VM_NOT)           \
  VM_CD(VM_SELECT) VM_CD(VM_LOADG)

static_assert(kVmMaxDepth == 10, "VM_CD covers depths 0..10");
static_assert(VM_NUM_OPS == 39, "VM_ALL_CASES lists every opcode");

template <int NCLS, int V, bool MATH>
__device__ __forceinline__ void vm_run(const VmProgram& P, double (&S)[kVmMaxDepth][V],
                                       const CellIdx<NCLS, V>& c)
{
  const int n = P.n_insns;
  for (int pc = 0; pc < n; ++pc) {
    const VmInsn in = P.insns[pc];
    const int imm = in.imm;
    switch (in.code) {
      VM_ALL_CASES
    default: break;
    }
  }
}

struct Box {
  int lo[3];
  int n[3];        // extent of the box per axis
  int nplane;      // n[0] * n[1]
  int zchunks;     // ceil(n[2] / V)
};

template <int NCLS, int V>
__device__ __forceinline__ void make_cell(const VmProgram& P, const Box& B, int t, int zc,
                                          CellIdx<NCLS, V>& c, int* nv)
{
  const int yy = t / B.n[0];
  c.x = B.lo[0] + (t - yy * B.n[0]);
  c.y = B.lo[1] + yy;
  const int z0 = B.lo[2] + zc * V;
  const int left = B.lo[2] + B.n[2] - z0;
  *nv = left < V ? left : V;
#pragma unroll
  for (int v = 0; v < V; ++v) c.z[v] = z0 + (v < *nv ? v : *nv - 1);  // clamp: stays in bounds
#pragma unroll
  for (int k = 0; k < NCLS; ++k) {
    const int X = P.cls_X[k], XY = P.cls_XY[k];
#pragma unroll
    for (int v = 0; v < V; ++v) c.i[k][v] = c.x + X * c.y + XY * c.z[v];
  }
}

template <int NCLS, int V, bool MATH>
__global__ __launch_bounds__(kBlock) void vm_assign_kernel(const VmProgram P, Box B,
                                                           double* __restrict__ out, int oX,
                                                           int oXY)
{
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= B.nplane) return;
  CellIdx<NCLS, V> c;
  int nv;
  make_cell<NCLS, V>(P, B, t, blockIdx.y, c, &nv);
  double S[kVmMaxDepth][V];
  vm_run<NCLS, V, MATH>(P, S, c);
  const int base = c.x + oX * c.y;
#pragma unroll
  for (int v = 0; v < V; ++v)
    if (v < nv) out[base + oXY * c.z[v]] = S[0][v];
}

template <int NCLS, bool MATH>
__global__ __launch_bounds__(kBlock) void vm_masked_kernel(const VmProgram P,
                                                           const int32_t* __restrict__ pts,
                                                           int npts, double* out, int oX,
                                                           int oXY)
{
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= npts) return;
  CellIdx<NCLS, 1> c;
  c.x = pts[3 * t];
  c.y = pts[3 * t + 1];
  c.z[0] = pts[3 * t + 2];
#pragma unroll
  for (int k = 0; k < NCLS; ++k) c.i[k][0] = c.x + P.cls_X[k] * c.y + P.cls_XY[k] * c.z[0];
  double S[kVmMaxDepth][1];
  vm_run<NCLS, 1, MATH>(P, S, c);
  out[c.x + oX * c.y + oXY * c.z[0]] = S[0][0];
}

// ---- reductions ----------------------------------------------------------
// max/min use the reference's select forms ((a < b) ? b : a and (a > b) ? b : a),
// which skip NaN operands the same way the serial fold does.  Sums carry a
// double-double (hi, lo) pair built from error-free TwoSum steps.

struct DD { double hi, lo; };

__device__ __forceinline__ DD dd_add(DD a, double b)
{
  const double s = a.hi + b;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b - bb);
  return DD{s, a.lo + e};
}

__device__ __forceinline__ DD dd_merge(DD a, DD b)
{
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double e = ((a.hi - (s - bb)) + (b.hi - bb)) + (a.lo + b.lo);
  const double hi = s + e;
  return DD{hi, e - (hi - s)};
}

__device__ __forceinline__ double red_max(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double red_min(double a, double b) { return (a > b) ? b : a; }

template <int OP>
__device__ __forceinline__ DD red_combine(DD a, DD b)
{
  if constexpr (OP == NEBO_RED_MAX) return DD{red_max(a.hi, b.hi), 0.0};
  else if constexpr (OP == NEBO_RED_MIN) return DD{red_min(a.hi, b.hi), 0.0};
  else return dd_merge(a, b);
}

template <int OP>
__device__ __forceinline__ DD red_init()
{
  if constexpr (OP == NEBO_RED_MAX) return DD{-INFINITY, 0.0};
  else if constexpr (OP == NEBO_RED_MIN) return DD{INFINITY, 0.0};
  else return DD{0.0, 0.0};
}

// Block-wide fixed-order reduction; returns the total in thread 0.
template <int OP>
__device__ DD block_reduce(DD v)
{
  __shared__ double sh_hi[kBlock / 64], sh_lo[kBlock / 64];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    DD o{__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)};
    v = red_combine<OP>(v, o);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sh_hi[wave] = v.hi; sh_lo[wave] = v.lo; }
  __syncthreads();
  if (threadIdx.x == 0) {
    DD r{sh_hi[0], sh_lo[0]};
    for (int w = 1; w < kBlock / 64; ++w) r = red_combine<OP>(r, DD{sh_hi[w], sh_lo[w]});
    v = r;
  }
  return v;
}

// The result of a reduction from its block partials: the partials folded in
// index order (thread t takes partials t, t + kBlock, ...) and then the block's
// fixed tree, so the value is deterministic; max/min return the selected element
// itself (a -0.0 stays -0.0).
template <int OP>
__device__ void fold_partials(const double* partial, int n, double* out)
{
  // The partials were just written by blocks on every XCD, so each load waits on
  // memory: a thread issues the loads of up to U of its partials before it
  // combines them, in the same order as one at a time.  A reduction over 258^3
  // points (2048 partials) took 1.1-1.4 us less wall time this way (max 24.2
  // against 25.4 us).
  constexpr int U = 8;
  DD acc = red_init<OP>();
  for (int base = threadIdx.x; base < n; base += U * kBlock) {
    double hi[U], lo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kBlock;
      hi[u] = i < n ? partial[2 * i] : 0.0;
      lo[u] = i < n ? partial[2 * i + 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * kBlock < n) acc = red_combine<OP>(acc, DD{hi[u], lo[u]});
  }
  DD r = block_reduce<OP>(acc);
  if (threadIdx.x == 0) {
    double v = (OP == NEBO_RED_MAX || OP == NEBO_RED_MIN) ? r.hi : r.hi + r.lo;
    if constexpr (OP == NEBO_RED_NORM) v = __builtin_sqrt(v);
    *out = v;
  }
}

template <int NCLS, int V, int OP, bool MATH>
__global__ __launch_bounds__(kBlock) void vm_reduce_kernel(const VmProgram P, Box B,
                                                           double* __restrict__ partial)
{
  DD acc = red_init<OP>();
  const long items = (long)B.nplane * B.zchunks;
  for (long w = (long)blockIdx.x * kBlock + threadIdx.x; w < items;
       w += (long)gridDim.x * kBlock) {
    const int zc = (int)(w / B.nplane);
    const int t = (int)(w - (long)zc * B.nplane);
    CellIdx<NCLS, V> c;
    int nv;
    make_cell<NCLS, V>(P, B, t, zc, c, &nv);
    double S[kVmMaxDepth][V];
    vm_run<NCLS, V, MATH>(P, S, c);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      if (v < nv) {
        const double x = S[0][v];
        if constexpr (OP == NEBO_RED_MAX) acc.hi = red_max(acc.hi, x);
        else if constexpr (OP == NEBO_RED_MIN) acc.hi = red_min(acc.hi, x);
        else if constexpr (OP == NEBO_RED_SUM) acc = dd_add(acc, x);
        else {
          // x*x exactly as p + err (Dekker split; no FMA so the result is portable)
          const double p = x * x;
          const double t2 = 134217729.0 * x;
          const double xh = t2 - (t2 - x), xl = x - xh;
          const double err = ((xh * xh - p) + 2.0 * xh * xl) + xl * xl;
          acc = dd_add(acc, p);
          acc.lo = acc.lo + err;
        }
      }
    }
  }
  DD r = block_reduce<OP>(acc);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r.hi;
    partial[2 * blockIdx.x + 1] = r.lo;
  }
}

// The second launch of every reduction: one block folds the n block partials
// into *out.  (Folding them in the main kernel's last block needs an agent-scope
// release per block, an L2 write-back, plus one same-address atomic per block
// across the XCDs: measured 74 us against 24 + 4 us for max over 258^3 points.)
template <int OP>
__global__ __launch_bounds__(kBlock) void reduce_final_kernel(const double* __restrict__ partial, int n,
                                                              double* __restrict__ out)
{
  fold_partials<OP>(partial, n, out);
}

Box make_box(const nebo_limits_t& L, int V)
{
  Box B;
  for (int d = 0; d < 3; ++d) { B.lo[d] = L.lo[d]; B.n[d] = L.hi[d] - L.lo[d]; }
  B.nplane = B.n[0] * B.n[1];
  B.zchunks = (B.n[2] + V - 1) / V;
  return B;
}

// Every field the program reads, and the LHS, must be addressable with 32-bit
// flat indices.
int check_index_range(const nebo_expr_t* e, const nebo_field_t* lhs, const char* ctx)
{
  auto big = [](const int32_t g[3]) { return (long)g[0] * g[1] * g[2] >= (1L << 31); };
  if (lhs && big(lhs->glob)) return set_error(NEBO_ERR_TOO_LARGE, ctx, "field has 2^31 or more points");
  for (int i = 0; i < e->n_fields; ++i)
    if (big(e->fields[i].glob)) return set_error(NEBO_ERR_TOO_LARGE, ctx, "field has 2^31 or more points");
  return NEBO_OK;
}

// Operands must live on the LHS device (NeboLhs.h:107-153 error texts).
int check_devices(const nebo_expr_t* e, int device, const char* ctx)
{
  if (device < 0)
    return set_error(NEBO_ERR_DEVICE, ctx,
                     "Left-hand side of assignment allocated on CPU; this engine evaluates on GPU "
                     "only (the CPU backend is the reference's SeqWalk loop)");
  for (int i = 0; i < e->n_fields; ++i)
    if (e->fields[i].device != device)
      return set_error(NEBO_ERR_DEVICE, ctx,
                       "Left-hand side of assignment allocated on GPU but right-hand side is not "
                       "(completely) accessible on the same GPU");
  return NEBO_OK;
}

bool subtree_reads(const nebo_expr_t* e, int node, const void* values)
{
  const nebo_node_t& n = e->nodes[node];
  switch (n.op) {
  case NEBO_SCALAR: case NEBO_MASK: case NEBO_MASK_SHIFT: return false;
  case NEBO_FIELD: return e->fields[n.idx].values == values;
  default:
    if (subtree_reads(e, n.a, values)) return true;
    if (n.op == NEBO_STENCIL || n.op == NEBO_NEG || n.op == NEBO_ABS || n.op == NEBO_SQRT ||
        n.op == NEBO_SQUARE || n.op == NEBO_CUBE || n.op == NEBO_POWINT || n.op == NEBO_NOT ||
        (n.op >= NEBO_EXP && n.op <= NEBO_ERFC) || n.op == NEBO_INV_ERF || n.op == NEBO_INV_ERFC)
      return false;
    if (subtree_reads(e, n.b, values)) return true;
    return n.op == NEBO_COND && subtree_reads(e, n.c, values);
  }
}

// An in-place assignment may read the LHS only at the point it writes.
int check_aliasing(const nebo_field_t* lhs, const nebo_expr_t* e, const char* ctx)
{
  for (int i = 0; i < e->n_nodes; ++i) {
    const nebo_node_t& s = e->nodes[i];
    if (s.op != NEBO_STENCIL) continue;
    if (subtree_reads(e, s.a, lhs->values))
      return set_error(NEBO_ERR_INVALID, ctx,
                       "left-hand side is read by a stencil on the right-hand side; neighbouring "
                       "points would be read after they are written (assign into another field)");
  }
  return NEBO_OK;
}

// Flat index of interior-relative point q in field f's allocation.
long flat_of(const nebo_field_t& f, const int32_t q[3])
{
  const long x = f.offset[0] + f.valid_minus[0] + q[0];
  const long y = f.offset[1] + f.valid_minus[1] + q[1];
  const long z = f.offset[2] + f.valid_minus[2] + q[2];
  return x + (long)f.glob[0] * (y + (long)f.glob[1] * z);
}

std::string point_text(const int32_t* p)
{
  std::ostringstream m;
  m << "(" << p[0] << ", " << p[1] << ", " << p[2] << ")";
  return m.str();
}

// Reads of a masked assignment, point by point: edgeless reads must stay in
// their field's block, and the LHS may only be read where no other point writes.
int check_masked_reads(const nebo_field_t* lhs, const nebo_expr_t* e, const int32_t* points,
                       int n_points, const char* ctx)
{
  std::vector<FieldRead> reads;
  collect_reads(e, e->n_nodes - 1, 0, 0, 0, false, &reads);
  bool reads_lhs = false;
  for (const FieldRead& r : reads) {
    const nebo_field_t& f = e->fields[r.field];
    if (f.values == lhs->values) reads_lhs = true;
    if (!r.edgeless) continue;
    for (int i = 0; i < n_points; ++i)
      for (int d = 0; d < 3; ++d) {
        const long g = (long)f.offset[d] + f.valid_minus[d] + points[3 * i + d] + r.s[d];
        if (g < 0 || g >= f.glob[d]) {
          std::ostringstream m;
          m << "mask point " << i << " " << point_text(points + 3 * i) << " reads field " << r.field
            << " at offset (" << r.s[0] << ", " << r.s[1] << ", " << r.s[2]
            << "), outside the field's allocated block";
          return set_error(NEBO_ERR_GHOST, ctx, m.str());
        }
      }
  }
  if (!reads_lhs) return NEBO_OK;
  std::unordered_map<long, int> written;
  written.reserve((size_t)n_points * 2);
  for (int i = 0; i < n_points; ++i)
    if (!written.emplace(flat_of(*lhs, points + 3 * i), i).second)
      return set_error(NEBO_ERR_INVALID, ctx,
                       "mask lists point " + point_text(points + 3 * i) +
                           " more than once while the right-hand side reads the left-hand side");
  for (const FieldRead& r : reads) {
    const nebo_field_t& f = e->fields[r.field];
    if (f.values != lhs->values) continue;
    for (int i = 0; i < n_points; ++i) {
      const int32_t q[3] = {points[3 * i] + r.s[0], points[3 * i + 1] + r.s[1],
                            points[3 * i + 2] + r.s[2]};
      auto it = written.find(flat_of(f, q));
      if (it != written.end() && it->second != i)
        return set_error(NEBO_ERR_INVALID, ctx,
                         "mask point " + point_text(points + 3 * i) + " reads the left-hand side at " +
                             point_text(q) + ", which mask point " +
                             point_text(points + 3 * it->second) +
                             " writes; the result would depend on the order of the writes");
    }
  }
  return NEBO_OK;
}

}  // namespace

bool uses_math(const VmProgram& P)
{
  for (int i = 0; i < P.n_insns; ++i) {
    const int op = P.insns[i].code >> 4;
    if (op == VM_POW || op == VM_POWC || (op >= VM_EXP && op <= VM_INV_ERFC)) return true;
  }
  return false;
}

int launch_generic_assign(const nebo_field_t* lhs, const VmProgram& prog, int max_depth,
                          const nebo_limits_t& lim, void* stream)
{
  (void)max_depth;
  bool done = false;
  int rc = jit_assign(lhs, prog, lim, stream, &done);
  if (rc || done) return rc;
  Box B = make_box(lim, kVmV);
  dim3 grid((B.nplane + kBlock - 1) / kBlock, B.zchunks);
  double* out = frame_origin(lhs);
  const int oX = lhs->glob[0], oXY = lhs->glob[0] * lhs->glob[1];
  const bool one = prog.n_classes == 1, math = uses_math(prog);
  BackendScope bs("stack-program interpreter (" + std::to_string(prog.n_insns) + " instructions)");
  return emit(stream, CallPtrs{}, [prog, B, grid, out, oX, oXY, one, math](void* sv, const CallPtrs&) {
    hipStream_t s = (hipStream_t)sv;
    if (one && !math)
      hipLaunchKernelGGL((vm_assign_kernel<1, kVmV, false>), grid, dim3(kBlock), 0, s, prog, B, out, oX, oXY);
    else if (one)
      hipLaunchKernelGGL((vm_assign_kernel<1, kVmV, true>), grid, dim3(kBlock), 0, s, prog, B, out, oX, oXY);
    else if (!math)
      hipLaunchKernelGGL((vm_assign_kernel<kVmMaxClasses, kVmV, false>), grid, dim3(kBlock), 0, s, prog, B, out, oX,
                         oXY);
    else
      hipLaunchKernelGGL((vm_assign_kernel<kVmMaxClasses, kVmV, true>), grid, dim3(kBlock), 0, s, prog, B, out, oX,
                         oXY);
    return check_hip(hipGetLastError(), "Nebo Assignment");
  });
}

namespace {

int plan_assign(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost, void* stream,
                bool allow_special);

// lhs <<= rhs: replays the launches of an earlier identical call, else checks,
// plans and launches the assignment, recording its launches (GPU operands).
int do_assign(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost, void* stream,
              bool allow_special)
{
  static const char* ctx = "Nebo Assignment";
  clear_error();
  if (!lhs || !rhs) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  thread_local CallKey key;
  int device = 0;
  bool keyed = lhs->device >= 0;
  if (keyed && hipGetDevice(&device) != hipSuccess) {
    (void)hipGetLastError();
    keyed = false;
  }
  if (keyed) {
    key.begin(allow_special ? 'A' : 'G', device, use_ghost, 0);
    key.field(*lhs);
    keyed = key.expr(rhs);
  }
  if (!keyed) return plan_assign(lhs, rhs, use_ghost, stream, allow_special);
  if (auto r = memo_find(key)) return memo_run(*r, stream, CallPtrs{});
  MemoRecorder rec(key);
  const int rc = plan_assign(lhs, rhs, use_ghost, stream, allow_special);
  rec.commit(rc);
  return rc;
}

int plan_assign(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost, void* stream,
                bool allow_special)
{
  static const char* ctx = "Nebo Assignment";
  nebo_limits_t lim;
  int empty = 0;
  int rc = assign_limits(lhs, rhs, use_ghost, &lim, &empty);
  if (rc) return rc;
  if ((rc = check_devices(rhs, lhs->device, ctx))) return rc;
  if ((rc = check_index_range(rhs, lhs, ctx))) return rc;
  if ((rc = check_aliasing(lhs, rhs, ctx))) return rc;
  if (empty) return NEBO_OK;
  if (allow_special) {
    int kid = K_GENERIC;
    if ((rc = nebo_assign_plan(lhs, rhs, use_ghost, &kid))) return rc;
    if (kid == K_DIVKGRAD) {
      DivKGradMatch m;
      match_divkgrad(rhs, &m);
      return launch_divkgrad(lhs, &rhs->fields[m.k_field], &rhs->fields[m.phi_field], m.c,
                             m.form, lim, stream);
    }
    if (kid == K_LAPL_EULER) {
      LaplEulerMatch m;
      match_lapl_euler(rhs, &m);
      return launch_lapl_euler(lhs, &rhs->fields[m.phi_field], m.c, m.s, lim, stream);
    }
  }
  VmProgram prog;
  int depth = 0;
  if ((rc = lower_program(rhs, &prog, &depth))) return rc;
  return launch_generic_assign(lhs, prog, depth, lim, stream);
}

template <int NCLS, int OP, bool MATH>
Launch reduce_op_launch(const VmProgram& prog, const Box& B, int blocks)
{
  return [prog, B, blocks](void* s, const CallPtrs& p) {
    hipLaunchKernelGGL((vm_reduce_kernel<NCLS, kVmV, OP, MATH>), dim3(blocks), dim3(kBlock), 0, (hipStream_t)s, prog,
                       B, p.partial);
    return check_hip(hipGetLastError(), "Nebo Reduction");
  };
}

template <int NCLS, bool MATH>
Launch reduce_any_launch(const VmProgram& prog, const Box& B, int op, int blocks)
{
  switch (op) {
  case NEBO_RED_MAX: return reduce_op_launch<NCLS, NEBO_RED_MAX, MATH>(prog, B, blocks);
  case NEBO_RED_MIN: return reduce_op_launch<NCLS, NEBO_RED_MIN, MATH>(prog, B, blocks);
  case NEBO_RED_SUM: return reduce_op_launch<NCLS, NEBO_RED_SUM, MATH>(prog, B, blocks);
  default: return reduce_op_launch<NCLS, NEBO_RED_NORM, MATH>(prog, B, blocks);
  }
}

// The fold of `blocks` partials (p.partial) into the result slot p.out.
Launch reduce_final_launch(int op, int blocks)
{
  return [op, blocks](void* sv, const CallPtrs& p) {
    hipStream_t s = (hipStream_t)sv;
    switch (op) {
    case NEBO_RED_MAX: hipLaunchKernelGGL((reduce_final_kernel<NEBO_RED_MAX>), dim3(1), dim3(kBlock), 0, s, p.partial, blocks, p.out); break;
    case NEBO_RED_MIN: hipLaunchKernelGGL((reduce_final_kernel<NEBO_RED_MIN>), dim3(1), dim3(kBlock), 0, s, p.partial, blocks, p.out); break;
    case NEBO_RED_SUM: hipLaunchKernelGGL((reduce_final_kernel<NEBO_RED_SUM>), dim3(1), dim3(kBlock), 0, s, p.partial, blocks, p.out); break;
    default: hipLaunchKernelGGL((reduce_final_kernel<NEBO_RED_NORM>), dim3(1), dim3(kBlock), 0, s, p.partial, blocks, p.out); break;
    }
    return check_hip(hipGetLastError(), "Nebo Reduction");
  };
}

// Two launches: the kernel compiled for the tree's shape (compiled now if
// `compile`, else only when a previous call compiled it) or the interpreter's
// kernel writes one partial per block into ptrs.partial; the final kernel folds
// them into ptrs.out.
int launch_reduce(const VmProgram& prog, const nebo_limits_t& lim, int op, const CallPtrs& ptrs, bool compile,
                  hipStream_t s)
{
  const Box B = make_box(lim, kVmV);
  const long items = (long)B.nplane * B.zchunks;
  int blocks = (int)((items + kBlock - 1) / kBlock);
  if (blocks > kRedMaxBlocks) blocks = kRedMaxBlocks;
  if (blocks < 1) blocks = 1;
  static_assert(kRedBufBytes >= sizeof(double) * (2 * kRedMaxBlocks + 1), "partials and result fit the buffer");
  bool done = false;
  int jit_blocks = 0;
  int rc = jit_reduce(prog, lim, op, ptrs, kRedMaxBlocks, compile, s, &jit_blocks, &done);
  if (rc) return rc;
  if (done) return emit(s, ptrs, reduce_final_launch(op, jit_blocks));
  const bool one = prog.n_classes == 1, math = uses_math(prog);
  {
    BackendScope bs("stack-program interpreter reduction (" + std::to_string(blocks) + " blocks)");
    Launch l = one && !math ? reduce_any_launch<1, false>(prog, B, op, blocks)
               : one        ? reduce_any_launch<1, true>(prog, B, op, blocks)
               : !math      ? reduce_any_launch<kVmMaxClasses, false>(prog, B, op, blocks)
                            : reduce_any_launch<kVmMaxClasses, true>(prog, B, op, blocks);
    if ((rc = emit(s, ptrs, std::move(l)))) return rc;
  }
  return emit(s, ptrs, reduce_final_launch(op, blocks));
}

// Reduction buffers: one per (device, stream), created by the first reduction
// on that stream and kept until the process ends.  Reductions on one stream run
// in stream order and share it; reductions on other streams use their own.
struct StreamRedBuf {
  RedBuf buf;
  std::mutex* mu;   // held by nebo_reduce from its launch to its copy of the result
};
std::mutex g_redbuf_mu;
std::map<std::pair<int, void*>, StreamRedBuf> g_redbuf;

int stream_reduce_buffer(void* stream, StreamRedBuf* out)
{
  static const char* ctx = "Nebo Reduction";
  int device = 0;
  int rc = check_hip(hipGetDevice(&device), ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_redbuf_mu);
  const auto key = std::make_pair(device, stream);
  auto it = g_redbuf.find(key);
  if (it != g_redbuf.end()) {
    *out = it->second;
    return NEBO_OK;
  }
  char* p = nullptr;
  if ((rc = check_hip(hipMalloc((void**)&p, kRedBufBytes), ctx))) return rc;
  StreamRedBuf b{red_buf_at(p), new std::mutex};
  g_redbuf[key] = b;
  *out = b;
  return NEBO_OK;
}

__global__ void fill_value_kernel(double* out, double v) { *out = v; }

// Reduction of e into ptrs.out (device memory) on stream, with ptrs.partial as
// working memory; `compile`: a tree shape met for the first time is compiled now
// (else it runs on the interpreter's kernel until an allocating call compiles it).
int do_reduce(const nebo_expr_t* e, int op, int use_ghost, const CallPtrs& ptrs, void* stream, bool compile)
{
  static const char* ctx = "Nebo Reduction";
  nebo_limits_t lim;
  int empty = 0;
  int rc = reduce_limits(e, use_ghost, &lim, &empty);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (empty) {
    const double v = op == NEBO_RED_MAX ? -INFINITY : op == NEBO_RED_MIN ? INFINITY : 0.0;
    return emit(s, ptrs, [v](void* sv, const CallPtrs& p) {
      hipLaunchKernelGGL(fill_value_kernel, dim3(1), dim3(1), 0, (hipStream_t)sv, p.out, v);
      return check_hip(hipGetLastError(), ctx);
    });
  }
  VmProgram prog;
  int depth = 0;
  if ((rc = lower_program(e, &prog, &depth))) return rc;
  return launch_reduce(prog, lim, op, ptrs, compile, s);
}

// do_reduce, its launches recorded under *key when key is non-null.
int recorded_reduce(const CallKey* key, const nebo_expr_t* e, int op, int use_ghost, const CallPtrs& ptrs,
                    void* stream, bool compile)
{
  if (!key) return do_reduce(e, op, use_ghost, ptrs, stream, compile);
  MemoRecorder rec(*key);
  const int rc = do_reduce(e, op, use_ghost, ptrs, stream, compile);
  rec.commit(rc);
  return rc;
}

// Key of a reduction; false (not recorded) for operands on the host or when no
// GPU is current.
bool reduce_key(CallKey* key, const nebo_expr_t* e, int op, int use_ghost, bool compile)
{
  if (!e || e->n_fields < 0 || (e->n_fields > 0 && (!e->fields || e->fields[0].device < 0))) return false;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  key->begin('R', device, op, 2 * use_ghost + (compile ? 1 : 0));
  return key->expr(e);
}

// Argument checks shared by the reduction entry points.
int check_reduce(const nebo_expr_t* e, int op, int use_ghost)
{
  static const char* ctx = "Nebo Reduction";
  if (!e) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  if (op < NEBO_RED_MAX || op > NEBO_RED_NORM) return set_error(NEBO_ERR_INVALID, ctx, "unknown reduction");
  nebo_limits_t lim;
  int empty = 0;
  int rc = reduce_limits(e, use_ghost, &lim, &empty);   // the tree's own errors first
  if (rc) return rc;
  int device = -1;
  for (int i = 0; i < e->n_fields; ++i) { device = e->fields[i].device; break; }
  if (e->n_fields > 0 && (rc = check_devices(e, device, ctx))) return rc;
  return check_index_range(e, nullptr, ctx);
}

}  // namespace
}  // namespace nebo

using namespace nebo;

extern "C" int nebo_assign(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                           void* stream)
{
  return do_assign(lhs, rhs, use_ghost, stream, true);
}

extern "C" int nebo_assign_generic(const nebo_field_t* lhs, const nebo_expr_t* rhs,
                                   int use_ghost, void* stream)
{
  return do_assign(lhs, rhs, use_ghost, stream, false);
}

extern "C" int nebo_masked_assign(const nebo_field_t* lhs, const int32_t* points,
                                  int32_t n_points, const nebo_expr_t* rhs, void* stream)
{
  static const char* ctx = "Nebo Masked Assignment";
  clear_error();
  if (!lhs || !rhs || (n_points > 0 && !points))
    return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  int rc = validate_expr(rhs, ctx);
  if (rc) return rc;
  if ((rc = check_devices(rhs, lhs->device, ctx))) return rc;
  if ((rc = check_index_range(rhs, lhs, ctx))) return rc;
  if (n_points <= 0) return NEBO_OK;
  // every point must lie in the LHS window and in the RHS's valid range
  Ghost g;
  if ((rc = node_ghosts(rhs, rhs->n_nodes - 1, true, &g))) return rc;
  for (int i = 0; i < n_points; ++i)
    for (int d = 0; d < 3; ++d) {
      const int p = points[3 * i + d];
      const int ext = lhs->extent[d] - lhs->valid_minus[d] - lhs->valid_plus[d];
      const int wlo = -lhs->valid_minus[d], whi = ext + lhs->valid_plus[d];
      const int rlo = g.infinite() ? wlo : -g.m[d], rhi = g.infinite() ? whi : ext + g.p[d];
      if (p < wlo || p >= whi || p < rlo || p >= rhi) {
        std::ostringstream m;
        m << "mask point " << i << " (" << points[3 * i] << ", " << points[3 * i + 1] << ", "
          << points[3 * i + 2] << ") lies outside the valid range of the assignment";
        return set_error(NEBO_ERR_GHOST, ctx, m.str());
      }
    }
  VmProgram prog;
  int depth = 0;
  if ((rc = lower_program(rhs, &prog, &depth))) return rc;
  if ((rc = check_masked_reads(lhs, rhs, points, n_points, ctx))) return rc;
  double* out = frame_origin(lhs);
  const int oX = lhs->glob[0], oXY = lhs->glob[0] * lhs->glob[1];
  hipStream_t s = (hipStream_t)stream;
  int32_t* pts = nullptr;
  const size_t bytes = sizeof(int32_t) * 3 * (size_t)n_points;
  if ((rc = check_hip(hipMallocAsync((void**)&pts, bytes, s), ctx))) return rc;
  if ((rc = check_hip(hipMemcpyAsync(pts, points, bytes, hipMemcpyHostToDevice, s), ctx))) return rc;
  dim3 grid((n_points + kBlock - 1) / kBlock);
  const bool one = prog.n_classes == 1, math = uses_math(prog);
  BackendScope bs("masked assignment (" + std::to_string(n_points) + " points)");
  if (one && !math)
    hipLaunchKernelGGL((vm_masked_kernel<1, false>), grid, dim3(kBlock), 0, s, prog, pts, n_points, out, oX, oXY);
  else if (one)
    hipLaunchKernelGGL((vm_masked_kernel<1, true>), grid, dim3(kBlock), 0, s, prog, pts, n_points, out, oX, oXY);
  else if (!math)
    hipLaunchKernelGGL((vm_masked_kernel<kVmMaxClasses, false>), grid, dim3(kBlock), 0, s, prog, pts, n_points, out, oX, oXY);
  else
    hipLaunchKernelGGL((vm_masked_kernel<kVmMaxClasses, true>), grid, dim3(kBlock), 0, s, prog, pts, n_points, out, oX, oXY);
  rc = check_hip(hipGetLastError(), ctx);
  (void)hipFreeAsync(pts, s);
  return rc;
}

extern "C" int nebo_reduce_async(const nebo_expr_t* expr, int op, int use_ghost,
                                 double* out_dev, void* stream)
{
  clear_error();
  thread_local CallKey key;
  const bool keyed = out_dev && reduce_key(&key, expr, op, use_ghost, true);
  int rc;
  if (keyed) {
    if (auto r = memo_find(key)) {
      StreamRedBuf b;
      if ((rc = stream_reduce_buffer(stream, &b))) return rc;
      return memo_run(*r, stream, CallPtrs{out_dev, b.buf.partial});
    }
  }
  if ((rc = check_reduce(expr, op, use_ghost))) return rc;
  if (!out_dev) return set_error(NEBO_ERR_INVALID, "Nebo Reduction", "null argument");
  StreamRedBuf b;
  if ((rc = stream_reduce_buffer(stream, &b))) return rc;
  return recorded_reduce(keyed ? &key : nullptr, expr, op, use_ghost, CallPtrs{out_dev, b.buf.partial}, stream,
                         true);
}

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per device
// (with that device current) and kept.
int scratch_on_device(int device, char** out)
{
  static const char* ctx = "Nebo Reduction";
  if (device < 0 || device >= kScratchDevices)
    return set_error(NEBO_ERR_DEVICE, ctx, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    char* p = nullptr;
    int rc = check_hip(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)), ctx);
    if (rc) return rc;
    if (!p) return set_error(NEBO_ERR_HIP, ctx, "no device address for the reduction scratch");
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return NEBO_OK;
}
}  // namespace

extern "C" int nebo_reduce_async_noalloc(const nebo_expr_t* expr, int op, int use_ghost, double* out_dev,
                                         void* stream)
{
  static const char* ctx = "Nebo Reduction";
  clear_error();
  thread_local CallKey key;
  const bool keyed = out_dev && reduce_key(&key, expr, op, use_ghost, false);
  int rc;
  int device = 0;
  char* p = nullptr;
  if (keyed) {
    if (auto r = memo_find(key)) {
      if ((rc = check_hip(hipGetDevice(&device), ctx))) return rc;
      if ((rc = scratch_on_device(device, &p))) return rc;
      return memo_run(*r, stream, CallPtrs{out_dev, red_buf_at(p).partial});
    }
  }
  if ((rc = check_reduce(expr, op, use_ghost))) return rc;
  if (!out_dev) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  if ((rc = check_hip(hipGetDevice(&device), ctx))) return rc;
  for (int i = 0; i < expr->n_fields; ++i)
    if (expr->fields[i].device != device)
      return set_error(NEBO_ERR_DEVICE, ctx, "operands must be resident on the current GPU");
  if ((rc = scratch_on_device(device, &p))) return rc;
  return recorded_reduce(keyed ? &key : nullptr, expr, op, use_ghost, CallPtrs{out_dev, red_buf_at(p).partial},
                         stream, false);
}

extern "C" int nebo_reduce(const nebo_expr_t* expr, int op, int use_ghost, double* out,
                           void* stream)
{
  static const char* ctx = "Nebo Reduction";
  clear_error();
  int rc = check_reduce(expr, op, use_ghost);
  if (rc) return rc;
  if (!out) return set_error(NEBO_ERR_INVALID, ctx, "null argument");
  hipStream_t s = (hipStream_t)stream;
  StreamRedBuf b;
  if ((rc = stream_reduce_buffer(stream, &b))) return rc;
  thread_local CallKey key;
  const bool keyed = reduce_key(&key, expr, op, use_ghost, true);
  {
    // the result slot is the stream's: no other nebo_reduce on this stream may
    // write it between this reduction and the copy of its result
    std::lock_guard<std::mutex> lk(*b.mu);
    const CallPtrs ptrs{b.buf.result, b.buf.partial};
    std::shared_ptr<const MemoRecord> r = keyed ? memo_find(key) : nullptr;
    rc = r ? memo_run(*r, stream, ptrs) : recorded_reduce(keyed ? &key : nullptr, expr, op, use_ghost, ptrs, stream, true);
    if (!rc) rc = check_hip(hipMemcpyAsync(out, b.buf.result, sizeof(double), hipMemcpyDeviceToHost, s), ctx);
  }
  if (!rc) rc = check_hip(hipStreamSynchronize(s), ctx);
  return rc;
}
