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
  int pad0;                                                                 \
  int clsX[4];                                                              \
  int clsXY[4];                                                             \
  const double* base[40];                                                   \
  int glo[40];                                                              \
  int ghi[40];                                                              \
  const unsigned* mbits[6];                                                 \
  int mo[6][3];                                                             \
  int mX[6];                                                                \
  int mY[6];                                                                \
  int pad1;                                                                 \
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

// Straight-line source for one program and V cells per thread.
std::string jit_source(const VmProgram& P, int V)
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
  std::ostringstream o;
  o << "struct JitArgs { " NEBO_JIT_STR(NEBO_JIT_ARGS_FIELDS) " };\n";
  o << "extern \"C\" __global__ __launch_bounds__(" << kJitBlock << ") void nebo_fused(const JitArgs A)\n{\n";
  o << "  const int t = (int)blockIdx.x * " << kJitBlock << " + (int)threadIdx.x;\n";
  o << "  if (t >= A.nplane) return;\n";
  o << "  const int yy = t / A.n0;\n  const int x = A.lo0 + (t - yy * A.n0);\n  const int y = A.lo1 + yy;\n";
  o << "  const int zb = A.lo2 + (int)blockIdx.y * " << V << ";\n";
  o << "  const int left = A.lo2 + A.n2 - zb;\n  const int nv = left < " << V << " ? left : " << V << ";\n";
  for (int v = 0; v < V; ++v) {
    // cells past the box end repeat the last cell (in bounds); only stores are guarded
    o << "  const int z" << v << " = zb + (" << v << " < nv ? " << v << " : nv - 1);\n";
    for (int k = 0; k < P.n_classes; ++k)
      o << "  const int i" << v << "_" << k << " = x + A.clsX[" << k << "] * y + A.clsXY[" << k << "] * z" << v
        << ";\n";
  }
  // phase 1: every read, for every cell
  for (int v = 0; v < V; ++v) {
    for (int s = 0; s < kVmMaxLoads; ++s) {
      if (!used[s]) continue;
      const std::string idx = "i" + std::to_string(v) + "_" + std::to_string(P.loads[s].cls);
      if (guarded[s])
        o << "  double L" << v << "_" << s << " = __builtin_nan(\"\");\n  if (" << idx << " >= A.glo[" << s
          << "] && " << idx << " < A.ghi[" << s << "]) L" << v << "_" << s << " = A.base[" << s << "][" << idx
          << "];\n";
      else
        o << "  const double L" << v << "_" << s << " = A.base[" << s << "][" << idx << "];\n";
    }
    for (int m = 0; m < kVmMaxMasks; ++m) {
      if (!mused[m]) continue;
      o << "  const int q" << v << "_" << m << " = (A.mo[" << m << "][0] + x) + A.mX[" << m << "] * ((A.mo[" << m
        << "][1] + y) + A.mY[" << m << "] * (A.mo[" << m << "][2] + z" << v << "));\n";
      o << "  const double M" << v << "_" << m << " = ((A.mbits[" << m << "][q" << v << "_" << m << " >> 5] >> (q"
        << v << "_" << m << " & 31)) & 1u) ? 1.0 : 0.0;\n";
    }
  }
  // phase 2: the program, one SSA temporary per result
  for (int v = 0; v < V; ++v) {
    std::vector<std::string> st;
    int tmp = 0;
    auto fresh = [&](const std::string& expr) {
      const std::string name = "t" + std::to_string(v) + "_" + std::to_string(tmp++);
      o << "  const double " << name << " = " << expr << ";\n";
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
    // a bare leaf has no temporary: give the result a name
    const std::string r = st.empty() ? std::string("0.0") : st.back();
    o << "  if (" << v << " < nv) __builtin_nontemporal_store(" << r << ", &A.out[x + A.oX * y + A.oXY * z" << v
      << "]);\n";
  }
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

int compile(const std::string& src, int device, Entry* e)
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
  if ((rc = check_hip(hipModuleGetFunction(&e->fn, e->mod, "nebo_fused"), ctx))) {
    (void)hipModuleUnload(e->mod);
    e->mod = nullptr;
    return rc;
  }
  return NEBO_OK;
}

}  // namespace

bool jit_enabled() { return g_mode.load() != 0; }

int jit_assign(const nebo_field_t* lhs, const VmProgram& P, const nebo_limits_t& lim, void* stream, bool* done)
{
  static const char* ctx = "Nebo Fused Kernel";
  *done = false;
  if (!jit_enabled()) return NEBO_OK;
  const int nr = reads_of(P);
  const int V = nr <= 6 ? 4 : nr <= 16 ? 2 : 1;
  const std::string src = jit_source(P, V);
  int device = 0;
  int rc = check_hip(hipGetDevice(&device), ctx);
  if (rc) return rc;
  const std::string key = std::to_string(device) + "\n" + src;
  Entry e;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      e = it->second;
      ++g_hits;
    } else {
      Entry fresh;
      fresh.V = V;
      if (compile(src, device, &fresh)) {
        ++g_failures;
        clear_error();
        g_cache[key] = fresh;   // fn == nullptr: remembered failure, the interpreter runs it
        return NEBO_OK;
      }
      g_cache[key] = fresh;
      e = fresh;
      ++g_compiled;
    }
  }
  if (!e.fn) return NEBO_OK;
  JitArgs A;
  std::memset(&A, 0, sizeof A);
  A.out = frame_origin(lhs);
  A.oX = lhs->glob[0];
  A.oXY = lhs->glob[0] * lhs->glob[1];
  A.lo0 = lim.lo[0];
  A.lo1 = lim.lo[1];
  A.lo2 = lim.lo[2];
  A.n0 = lim.hi[0] - lim.lo[0];
  A.n1 = lim.hi[1] - lim.lo[1];
  A.n2 = lim.hi[2] - lim.lo[2];
  A.nplane = A.n0 * A.n1;
  for (int k = 0; k < kVmMaxClasses; ++k) {
    A.clsX[k] = P.cls_X[k];
    A.clsXY[k] = P.cls_XY[k];
  }
  for (int s = 0; s < kVmMaxLoads; ++s) {
    A.base[s] = P.loads[s].base;
    A.glo[s] = P.loads[s].glo;
    A.ghi[s] = P.loads[s].ghi;
  }
  for (int m = 0; m < kVmMaxMasks; ++m) {
    A.mbits[m] = P.masks[m].bits;
    for (int d = 0; d < 3; ++d) A.mo[m][d] = P.masks[m].o[d];
    A.mX[m] = P.masks[m].X;
    A.mY[m] = P.masks[m].Y;
  }
  for (int c = 0; c < kVmMaxConsts; ++c) A.c[c] = P.consts[c];
  size_t bytes = sizeof A;
  void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &bytes, HIP_LAUNCH_PARAM_END};
  const unsigned gx = (unsigned)((A.nplane + kJitBlock - 1) / kJitBlock);
  const unsigned gy = (unsigned)((A.n2 + e.V - 1) / e.V);
  rc = check_hip(hipModuleLaunchKernel(e.fn, gx, gy, 1, kJitBlock, 1, 1, 0, (hipStream_t)stream, nullptr, cfg),
                 ctx);
  if (!rc) *done = true;
  return rc;
}

std::string jit_source_for(const VmProgram& P)
{
  const int nr = reads_of(P);
  return jit_source(P, nr <= 6 ? 4 : nr <= 16 ? 2 : 1);
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

extern "C" int nebo_fused_source(const nebo_expr_t* rhs, char* buf, int32_t size, int32_t* needed)
{
  nebo::clear_error();
  int rc = nebo::validate_expr(rhs, "Nebo Fused Kernel");
  if (rc) return rc;
  nebo::VmProgram prog;
  int depth = 0;
  if ((rc = nebo::lower_program(rhs, &prog, &depth))) return rc;
  const std::string s = nebo::jit_source_for(prog);
  if (needed) *needed = (int32_t)s.size() + 1;
  if (buf && size > 0) {
    const size_t n = s.size() < (size_t)size - 1 ? s.size() : (size_t)size - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return NEBO_OK;
}
