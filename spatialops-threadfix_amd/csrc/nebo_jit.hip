// nebo_jit.hip -- fused kernels compiled at run time for the trees that have no
// hand-written kernel.
//
// The stack program of nebo_lower.cpp is turned into straight-line HIP source:
// every distinct (field, offset) read and mask read is issued first, for V cells
// of a z column per thread, then the program's operations follow as SSA
// temporaries in program order (the order the reference's recursive eval uses,
// so results are bit-identical to the interpreter in nebo_eval.hip), then the
// nontemporal stores.  hiprtc compiles the source for gfx950 with the library's
// own flags (-ffp-contract=off, no packed fp32); the code object is loaded once
// per device and cached by source text.  Pointers, constants, index classes and
// the box are kernel arguments, so every tree of one shape shares one kernel.
// Constants stay arguments on purpose: a literal exponent would let the
// compiler rewrite pow(x, 2.0), which the interpreter evaluates through the
// math library.
//
// The argument struct is handed to hipModuleLaunchKernel as the kernel's
// argument buffer (HIP_LAUNCH_PARAM_BUFFER_POINTER), which is passed to the
// kernel as it is: the hidden arguments that follow the explicit ones in the
// kernarg segment (block counts behind gridDim, group sizes behind blockDim,
// global offsets) are not filled in.  Generated code therefore reads none of
// gridDim, blockDim or the global offsets; the grid size a grid-stride loop
// needs travels in JitArgs::nblocks.  Only blockIdx and threadIdx (hardware
// registers) are used.
//
// If hiprtc is unavailable or a compile fails, nebo_assign runs the
// interpreter instead (also a GPU kernel) and the failure is counted.

#include "nebo_internal.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace nebo {
namespace {

#define NEBO_JIT_ARGS_FIELDS                                                 \
  double* out;                                                              \
  int oX;                                                                   \
  int oXY;                                                                  \
  int lo0, lo1, lo2, n0, n1, n2;                                            \
  int nplane;                                                               \
  int nzc;                                                                  \
  int clsX[4];                                                              \
  int clsXY[4];                                                             \
  const double* base[40];                                                   \
  int glo[40];                                                              \
  int ghi[40];                                                              \
  const unsigned* mbits[6];                                                 \
  int mo[6][3];                                                             \
  int mX[6];                                                                \
  int mY[6];                                                                \
  int nblocks;                                                              \
  double c[48];

struct JitArgs {
  NEBO_JIT_ARGS_FIELDS
};
static_assert(kVmMaxLoads == 40 && kVmMaxMasks == 6 && kVmMaxConsts == 48 && kVmMaxClasses == 4,
              "JitArgs mirrors the program limits");

#define NEBO_JIT_STR2(...) #__VA_ARGS__
#define NEBO_JIT_STR(...) NEBO_JIT_STR2(__VA_ARGS__)

constexpr int kJitBlock = 256;

std::atomic<int> g_mode{1};
std::atomic<int> g_compiled{0}, g_hits{0}, g_failures{0};

struct Entry {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  int V = 1;
};
std::mutex g_mu;
std::unordered_map<std::string, Entry> g_cache;   // key: device + source

// Index set-up, reads and operations for V cells of one z column at (x, y,
// z0 .. z{V-1}); the generated code expects x, y and the z variables in scope
// and returns the name of each cell's value.
std::vector<std::string> emit_cells(std::ostringstream& o, const VmProgram& P, int V, const std::string& ind)
{
  // which slots are read, and whether every read of a slot is bounds-checked
  bool used[kVmMaxLoads] = {}, guarded[kVmMaxLoads], mused[kVmMaxMasks] = {};
  for (int s = 0; s < kVmMaxLoads; ++s) guarded[s] = true;
  for (int pc = 0; pc < P.n_insns; ++pc) {
    const int op = P.insns[pc].code >> 4, imm = P.insns[pc].imm;
    if (op == VM_LOAD) { used[imm] = true; guarded[imm] = false; }
    if (op == VM_LOADG) used[imm] = true;
    if (op == VM_MASK) mused[imm] = true;
  }
  for (int v = 0; v < V; ++v)
    for (int k = 0; k < P.n_classes; ++k)
      o << ind << "const int i" << v << "_" << k << " = x + A.clsX[" << k << "] * y + A.clsXY[" << k << "] * z"
        << v << ";\n";
  // phase 1: every read, for every cell
  for (int v = 0; v < V; ++v) {
    for (int s = 0; s < kVmMaxLoads; ++s) {
      if (!used[s]) continue;
      const std::string idx = "i" + std::to_string(v) + "_" + std::to_string(P.loads[s].cls);
      if (guarded[s])
        o << ind << "double L" << v << "_" << s << " = __builtin_nan(\"\");\n" << ind << "if (" << idx
          << " >= A.glo[" << s << "] && " << idx << " < A.ghi[" << s << "]) L" << v << "_" << s << " = A.base["
          << s << "][" << idx << "];\n";
      else
        o << ind << "const double L" << v << "_" << s << " = A.base[" << s << "][" << idx << "];\n";
    }
    for (int m = 0; m < kVmMaxMasks; ++m) {
      if (!mused[m]) continue;
      o << ind << "const int q" << v << "_" << m << " = (A.mo[" << m << "][0] + x) + A.mX[" << m
        << "] * ((A.mo[" << m << "][1] + y) + A.mY[" << m << "] * (A.mo[" << m << "][2] + z" << v << "));\n";
      o << ind << "const double M" << v << "_" << m << " = ((A.mbits[" << m << "][q" << v << "_" << m
        << " >> 5] >> (q" << v << "_" << m << " & 31)) & 1u) ? 1.0 : 0.0;\n";
    }
  }
  // phase 2: the program, one SSA temporary per result
  std::vector<std::string> results;
  for (int v = 0; v < V; ++v) {
    std::vector<std::string> st;
    int tmp = 0;
    auto fresh = [&](const std::string& expr) {
      const std::string name = "t" + std::to_string(v) + "_" + std::to_string(tmp++);
      o << ind << "const double " << name << " = " << expr << ";\n";
      st.push_back(name);
    };
    auto pop = [&]() {
      std::string s = st.back();
      st.pop_back();
      return s;
    };
    const std::string vs = std::to_string(v);
    for (int pc = 0; pc < P.n_insns; ++pc) {
      const int op = P.insns[pc].code >> 4, imm = P.insns[pc].imm;
      const std::string k = "A.c[" + std::to_string(imm) + "]";
      switch (op) {
      case VM_LOAD: case VM_LOADG: st.push_back("L" + vs + "_" + std::to_string(imm)); break;
      case VM_MASK: st.push_back("M" + vs + "_" + std::to_string(imm)); break;
      case VM_CONST: fresh(k); break;
      case VM_MULC: { const std::string a = pop(); fresh(a + " * " + k); break; }
      case VM_SELECT: {
        const std::string c = pop(), b = pop(), a = pop();
        fresh("(" + a + " != 0.0) ? " + b + " : " + c);
        break;
      }
      default: {
        if (op >= VM_ADD && op <= VM_OR) {
          const std::string b = pop(), a = pop();
          switch (op) {
          case VM_ADD: fresh(a + " + " + b); break;
          case VM_SUB: fresh(a + " - " + b); break;
          case VM_MUL: fresh(a + " * " + b); break;
          case VM_DIV: fresh(a + " / " + b); break;
          case VM_MAX: fresh("(" + a + " > " + b + ") ? " + a + " : " + b); break;
          case VM_MIN: fresh("(" + a + " < " + b + ") ? " + a + " : " + b); break;
          case VM_POW: fresh("pow(" + a + ", " + b + ")"); break;
          case VM_EQ: fresh("(" + a + " == " + b + ") ? 1.0 : 0.0"); break;
          case VM_NE: fresh("(" + a + " != " + b + ") ? 1.0 : 0.0"); break;
          case VM_LT: fresh("(" + a + " < " + b + ") ? 1.0 : 0.0"); break;
          case VM_LE: fresh("(" + a + " <= " + b + ") ? 1.0 : 0.0"); break;
          case VM_GT: fresh("(" + a + " > " + b + ") ? 1.0 : 0.0"); break;
          case VM_GE: fresh("(" + a + " >= " + b + ") ? 1.0 : 0.0"); break;
          case VM_AND: fresh("(" + a + " != 0.0 && " + b + " != 0.0) ? 1.0 : 0.0"); break;
          default: fresh("(" + a + " != 0.0 || " + b + " != 0.0) ? 1.0 : 0.0"); break;   // VM_OR
          }
        } else {
          const std::string a = pop();
          switch (op) {
          case VM_NEG: fresh("-" + a); break;
          case VM_ABS: fresh("fabs(" + a + ")"); break;
          case VM_SQRT: fresh("__builtin_sqrt(" + a + ")"); break;
          case VM_SQUARE: fresh(a + " * " + a); break;
          case VM_CUBE: fresh(a + " * " + a + " * " + a); break;
          case VM_POWC: fresh("pow(" + a + ", " + k + ")"); break;
          case VM_EXP: fresh("exp(" + a + ")"); break;
          case VM_LOG: fresh("log(" + a + ")"); break;
          case VM_LOG10: fresh("log10(" + a + ")"); break;
          case VM_SIN: fresh("sin(" + a + ")"); break;
          case VM_COS: fresh("cos(" + a + ")"); break;
          case VM_TAN: fresh("tan(" + a + ")"); break;
          case VM_TANH: fresh("tanh(" + a + ")"); break;
          case VM_ERF: fresh("erf(" + a + ")"); break;
          case VM_ERFC: fresh("erfc(" + a + ")"); break;
          default: fresh("((" + a + ") != 0.0) ? 0.0 : 1.0"); break;   // VM_NOT
          }
        }
      }
      }
    }
    // a bare leaf has no temporary of its own; the leaf's name is the value
    results.push_back(st.empty() ? std::string("0.0") : st.back());
  }
  return results;
}

// Assignment kernel: one thread per (x, y) point of the box and V z cells.
std::string jit_source(const VmProgram& P, int V)
{
  std::ostringstream o;
  o << "struct JitArgs { " NEBO_JIT_STR(NEBO_JIT_ARGS_FIELDS) " };\n";
  o << "extern \"C\" __global__ __launch_bounds__(" << kJitBlock << ") void nebo_fused(const JitArgs A)\n{\n";
  o << "  const int t = (int)blockIdx.x * " << kJitBlock << " + (int)threadIdx.x;\n";
  o << "  if (t >= A.nplane) return;\n";
  o << "  const int yy = t / A.n0;\n  const int x = A.lo0 + (t - yy * A.n0);\n  const int y = A.lo1 + yy;\n";
  o << "  const int zb = A.lo2 + (int)blockIdx.y * " << V << ";\n";
  o << "  const int left = A.lo2 + A.n2 - zb;\n  const int nv = left < " << V << " ? left : " << V << ";\n";
  // cells past the box end repeat the last cell (in bounds); only stores are guarded
  for (int v = 0; v < V; ++v)
    o << "  const int z" << v << " = zb + (" << v << " < nv ? " << v << " : nv - 1);\n";
  const std::vector<std::string> r = emit_cells(o, P, V, "  ");
  for (int v = 0; v < V; ++v)
    o << "  if (" << v << " < nv) __builtin_nontemporal_store(" << r[v] << ", &A.out[x + A.oX * y + A.oXY * z" << v
      << "]);\n";
  o << "}\n";
  return o.str();
}

// Reduction kernel: a grid-stride loop over (x, y, z chunk) items, V z cells
// per item, each thread folding its cells into a double-double (sum, norm) or
// a select-form max/min exactly as the interpreter's vm_reduce_kernel does;
// then a fixed-order block reduction writes one (hi, lo) partial per block to
// A.out, which reduce_final_kernel folds.
std::string jit_reduce_source(const VmProgram& P, int V, int op)
{
  std::ostringstream o;
  o << "struct JitArgs { " NEBO_JIT_STR(NEBO_JIT_ARGS_FIELDS) " };\n";
  const char* init = op == NEBO_RED_MAX ? "-__builtin_inf()" : op == NEBO_RED_MIN ? "__builtin_inf()" : "0.0";
  const bool dd = op == NEBO_RED_SUM || op == NEBO_RED_NORM;
  // combine (hi, lo) with (h2, l2) into (hi, lo): TwoSum merge, or the select forms
  std::string combine;
  if (op == NEBO_RED_MAX) combine = "hi = (hi < h2) ? h2 : hi;";
  else if (op == NEBO_RED_MIN) combine = "hi = (hi > h2) ? h2 : hi;";
  else
    combine = "{ const double s = hi + h2; const double bb = s - hi; "
              "const double e = ((hi - (s - bb)) + (h2 - bb)) + (lo + l2); "
              "const double h = s + e; lo = e - (h - s); hi = h; }";
  o << "extern \"C\" __global__ __launch_bounds__(" << kJitBlock << ") void nebo_reduce(const JitArgs A)\n{\n";
  o << "  double hi = " << init << ", lo = 0.0;\n";
  o << "  const long items = (long)A.nplane * A.nzc;\n";
  o << "  for (long w = (long)blockIdx.x * " << kJitBlock << " + threadIdx.x; w < items; w += (long)A.nblocks * "
    << kJitBlock << ") {\n";
  o << "    const int zc = (int)(w / A.nplane);\n    const int t = (int)(w - (long)zc * A.nplane);\n";
  o << "    const int yy = t / A.n0;\n    const int x = A.lo0 + (t - yy * A.n0);\n    const int y = A.lo1 + yy;\n";
  o << "    const int zb = A.lo2 + zc * " << V << ";\n";
  o << "    const int left = A.lo2 + A.n2 - zb;\n    const int nv = left < " << V << " ? left : " << V << ";\n";
  for (int v = 0; v < V; ++v)
    o << "    const int z" << v << " = zb + (" << v << " < nv ? " << v << " : nv - 1);\n";
  const std::vector<std::string> r = emit_cells(o, P, V, "    ");
  for (int v = 0; v < V; ++v) {
    o << "    if (" << v << " < nv) {\n      const double xv = " << r[v] << ";\n";
    if (op == NEBO_RED_MAX) o << "      hi = (hi < xv) ? xv : hi;\n";
    else if (op == NEBO_RED_MIN) o << "      hi = (hi > xv) ? xv : hi;\n";
    else if (op == NEBO_RED_SUM)
      o << "      const double s = hi + xv; const double bb = s - hi;\n"
           "      lo = lo + ((hi - (s - bb)) + (xv - bb)); hi = s;\n";
    else
      // x*x exactly as p + err (Dekker split, no FMA)
      o << "      const double p = xv * xv; const double t2 = 134217729.0 * xv;\n"
           "      const double xh = t2 - (t2 - xv), xl = xv - xh;\n"
           "      const double err = ((xh * xh - p) + 2.0 * xh * xl) + xl * xl;\n"
           "      const double s = hi + p; const double bb = s - hi;\n"
           "      lo = lo + ((hi - (s - bb)) + (p - bb)); hi = s; lo = lo + err;\n";
    o << "    }\n";
  }
  o << "  }\n";
  // fixed-order block reduction: xor-shuffle tree inside each wave, then the waves in order
  o << "  for (int off = 32; off >= 1; off >>= 1) {\n";
  o << "    const double h2 = __shfl_xor(hi, off, 64);\n";
  o << "    const double l2 = " << (dd ? "__shfl_xor(lo, off, 64)" : "0.0") << ";\n";
  o << "    " << combine << "\n  }\n";
  o << "  __shared__ double sh[" << 2 * (kJitBlock / 64) << "];\n";
  o << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n";
  o << "  if (lane == 0) { sh[2 * wave] = hi; sh[2 * wave + 1] = lo; }\n";
  o << "  __syncthreads();\n";
  o << "  if (threadIdx.x == 0) {\n    hi = sh[0]; lo = sh[1];\n";
  o << "    for (int w = 1; w < " << kJitBlock / 64 << "; ++w) {\n";
  o << "      const double h2 = sh[2 * w], l2 = sh[2 * w + 1];\n      " << combine << "\n    }\n";
  o << "    A.out[2 * blockIdx.x] = hi;\n    A.out[2 * blockIdx.x + 1] = " << (dd ? "lo" : "0.0") << ";\n  }\n";
  o << "}\n";
  return o.str();
}

int reads_of(const VmProgram& P)
{
  bool used[kVmMaxLoads + kVmMaxMasks] = {};
  int n = 0;
  for (int pc = 0; pc < P.n_insns; ++pc) {
    const int op = P.insns[pc].code >> 4, imm = P.insns[pc].imm;
    int slot = -1;
    if (op == VM_LOAD || op == VM_LOADG) slot = imm;
    if (op == VM_MASK) slot = kVmMaxLoads + imm;
    if (slot >= 0 && !used[slot]) { used[slot] = true; ++n; }
  }
  return n;
}

int compile(const std::string& src, const char* name, int device, Entry* e)
{
  static const char* ctx = "Nebo Fused Kernel";
  hipDeviceProp_t prop;
  int rc = check_hip(hipGetDeviceProperties(&prop, device), ctx);
  if (rc) return rc;
  std::string arch = prop.gcnArchName;
  arch = "--offload-arch=" + arch.substr(0, arch.find(':'));
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "nebo_fused.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return set_error(NEBO_ERR_HIP, ctx, "hiprtcCreateProgram failed");
  const char* opts[] = {arch.c_str(), "-O3", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return set_error(NEBO_ERR_HIP, ctx, std::string("compile failed: ") + hiprtcGetErrorString(r) + "\n" + log);
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  std::vector<char> code(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  if ((rc = check_hip(hipModuleLoadData(&e->mod, code.data()), ctx))) return rc;
  if ((rc = check_hip(hipModuleGetFunction(&e->fn, e->mod, name), ctx))) {
    (void)hipModuleUnload(e->mod);
    e->mod = nullptr;
    return rc;
  }
  return NEBO_OK;
}

// Compiled kernel for src, from the cache or compiled now; fn stays null when
// compilation failed (remembered, so the interpreter runs such trees without
// another attempt).
int get_kernel(const std::string& src, const char* name, int V, Entry* out)
{
  static const char* ctx = "Nebo Fused Kernel";
  int device = 0;
  int rc = check_hip(hipGetDevice(&device), ctx);
  if (rc) return rc;
  const std::string key = std::to_string(device) + "\n" + src;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    *out = it->second;
    ++g_hits;
    return NEBO_OK;
  }
  Entry fresh;
  fresh.V = V;
  if (compile(src, name, device, &fresh)) {
    ++g_failures;
    clear_error();
    fresh = Entry{};
  } else {
    ++g_compiled;
  }
  g_cache[key] = fresh;
  *out = fresh;
  return NEBO_OK;
}

int cells_per_thread(const VmProgram& P)
{
  const int nr = reads_of(P);
  return nr <= 6 ? 4 : nr <= 16 ? 2 : 1;
}

void fill_args(const VmProgram& P, const nebo_limits_t& lim, int V, JitArgs* A)
{
  std::memset(A, 0, sizeof *A);
  A->lo0 = lim.lo[0];
  A->lo1 = lim.lo[1];
  A->lo2 = lim.lo[2];
  A->n0 = lim.hi[0] - lim.lo[0];
  A->n1 = lim.hi[1] - lim.lo[1];
  A->n2 = lim.hi[2] - lim.lo[2];
  A->nplane = A->n0 * A->n1;
  A->nzc = (A->n2 + V - 1) / V;
  for (int k = 0; k < kVmMaxClasses; ++k) {
    A->clsX[k] = P.cls_X[k];
    A->clsXY[k] = P.cls_XY[k];
  }
  for (int s = 0; s < kVmMaxLoads; ++s) {
    A->base[s] = P.loads[s].base;
    A->glo[s] = P.loads[s].glo;
    A->ghi[s] = P.loads[s].ghi;
  }
  for (int m = 0; m < kVmMaxMasks; ++m) {
    A->mbits[m] = P.masks[m].bits;
    for (int d = 0; d < 3; ++d) A->mo[m][d] = P.masks[m].o[d];
    A->mX[m] = P.masks[m].X;
    A->mY[m] = P.masks[m].Y;
  }
  for (int c = 0; c < kVmMaxConsts; ++c) A->c[c] = P.consts[c];
}

int launch(const Entry& e, JitArgs* A, unsigned gx, unsigned gy, void* stream)
{
  size_t bytes = sizeof *A;
  void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &bytes, HIP_LAUNCH_PARAM_END};
  return check_hip(hipModuleLaunchKernel(e.fn, gx, gy, 1, kJitBlock, 1, 1, 0, (hipStream_t)stream, nullptr, cfg),
                   "Nebo Fused Kernel");
}

}  // namespace

bool jit_enabled() { return g_mode.load() != 0; }

int jit_assign(const nebo_field_t* lhs, const VmProgram& P, const nebo_limits_t& lim, void* stream, bool* done)
{
  *done = false;
  if (!jit_enabled()) return NEBO_OK;
  const int V = cells_per_thread(P);
  Entry e;
  int rc = get_kernel(jit_source(P, V), "nebo_fused", V, &e);
  if (rc || !e.fn) return rc;
  JitArgs A;
  fill_args(P, lim, V, &A);
  A.out = frame_origin(lhs);
  A.oX = lhs->glob[0];
  A.oXY = lhs->glob[0] * lhs->glob[1];
  const unsigned gx = (unsigned)((A.nplane + kJitBlock - 1) / kJitBlock);
  if ((rc = launch(e, &A, gx, (unsigned)A.nzc, stream))) return rc;
  *done = true;
  return NEBO_OK;
}

int jit_reduce(const VmProgram& P, const nebo_limits_t& lim, int op, double* partial, int max_blocks, void* stream,
               int* blocks, bool* done)
{
  *done = false;
  if (!jit_enabled()) return NEBO_OK;
  const int V = cells_per_thread(P);
  Entry e;
  int rc = get_kernel(jit_reduce_source(P, V, op), "nebo_reduce", V, &e);
  if (rc || !e.fn) return rc;
  JitArgs A;
  fill_args(P, lim, V, &A);
  A.out = partial;
  const long items = (long)A.nplane * A.nzc;
  long nb = (items + kJitBlock - 1) / kJitBlock;
  if (nb > max_blocks) nb = max_blocks;
  if (nb < 1) nb = 1;
  A.nblocks = (int)nb;
  if ((rc = launch(e, &A, (unsigned)nb, 1, stream))) return rc;
  *blocks = (int)nb;
  *done = true;
  return NEBO_OK;
}

std::string jit_source_for(const VmProgram& P, int reduce_op)
{
  const int V = cells_per_thread(P);
  return reduce_op < 0 ? jit_source(P, V) : jit_reduce_source(P, V, reduce_op);
}

}  // namespace nebo

extern "C" int nebo_set_fused_jit(int enable)
{
  nebo::clear_error();
  nebo::g_mode.store(enable ? 1 : 0);
  return NEBO_OK;
}

extern "C" int nebo_fused_jit_stats(int32_t* compiled, int32_t* cache_hits, int32_t* failures)
{
  nebo::clear_error();
  if (compiled) *compiled = nebo::g_compiled.load();
  if (cache_hits) *cache_hits = nebo::g_hits.load();
  if (failures) *failures = nebo::g_failures.load();
  return NEBO_OK;
}

extern "C" int nebo_fused_source(const nebo_expr_t* rhs, int32_t reduce_op, char* buf, int32_t size,
                                 int32_t* needed)
{
  nebo::clear_error();
  if (reduce_op != -1 && (reduce_op < NEBO_RED_MAX || reduce_op > NEBO_RED_NORM))
    return nebo::set_error(NEBO_ERR_INVALID, "Nebo Fused Kernel", "unknown reduction");
  int rc = nebo::validate_expr(rhs, "Nebo Fused Kernel");
  if (rc) return rc;
  nebo::VmProgram prog;
  int depth = 0;
  if ((rc = nebo::lower_program(rhs, &prog, &depth))) return rc;
  const std::string s = nebo::jit_source_for(prog, reduce_op);
  if (needed) *needed = (int32_t)s.size() + 1;
  if (buf && size > 0) {
    const size_t n = s.size() < (size_t)size - 1 ? s.size() : (size_t)size - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return NEBO_OK;
}
