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

#include <algorithm>
#include <atomic>
#include <functional>
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
  int ntx;                                                                  \
  double c[48];

struct JitArgs {
  NEBO_JIT_ARGS_FIELDS
};
static_assert(kVmMaxLoads == 40 && kVmMaxMasks == 6 && kVmMaxConsts == 48 && kVmMaxClasses == 4,
              "JitArgs mirrors the program limits");

#define NEBO_JIT_STR2(...) #__VA_ARGS__
#define NEBO_JIT_STR(...) NEBO_JIT_STR2(__VA_ARGS__)

constexpr int kReduceBlock = 256;   // threads per block of the reduction kernels

// Launch shape of the assignment kernels and the cells-per-thread rule (see
// choose_cells); 0 in nebo_fused_jit_tune restores these defaults.
// Defaults from a sweep at 256^3 (pointwise, cond, Laplacian and div(k grad phi)
// trees): 512-thread blocks over the flattened plane, XCD renumbering on, up to
// 4 cells per thread within 48 distinct reads.  Larger blocks let a block's rows
// share their y-neighbour rows in L1; 2-D tiles and 8 cells per thread measured
// slower.
constexpr int kDefaultBlock = 512, kDefaultTileX = 0, kDefaultMaxCells = 4, kDefaultLoadBudget = 48,
              kDefaultXcd = 1;
std::atomic<int> g_block{kDefaultBlock}, g_tile_x{kDefaultTileX}, g_max_cells{kDefaultMaxCells},
    g_load_budget{kDefaultLoadBudget}, g_xcd{kDefaultXcd};
constexpr int kXcds = 8;   // MI355X: 8 accelerator complex dies, each with its own L2

// Launch shape of an assignment kernel: threads per block; tile_x == 0 lays the
// block over consecutive points of the flattened x-y plane, else over a
// tile_x by block/tile_x tile of it; xcd renumbers blocks so that consecutive
// tiles (which share halo rows and planes) run on one XCD and share its L2.
struct Shape {
  int block, tile_x, xcd;
};

std::atomic<int> g_mode{1};
std::atomic<int> g_compiled{0}, g_hits{0}, g_failures{0};

struct Entry {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};
std::mutex g_mu;
std::unordered_map<std::string, Entry> g_cache;   // key: device + source

// Which load slots a program reads, whether every read of a slot is
// bounds-checked (VM_LOADG only), and which masks it reads.
struct SlotUse {
  bool used[kVmMaxLoads] = {}, guarded[kVmMaxLoads] = {}, mused[kVmMaxMasks] = {};
  // z chains of unguarded slots: chain[c] lists slots whose bases lie one plane
  // (cls_XY doubles) apart, in increasing z; cell v + 1's read of chain[c][j]
  // is then cell v's read of chain[c][j + 1], the same address.
  std::vector<std::vector<int>> chains;
};

SlotUse slot_use(const VmProgram& P)
{
  SlotUse u;
  for (int s = 0; s < kVmMaxLoads; ++s) u.guarded[s] = true;
  for (int pc = 0; pc < P.n_insns; ++pc) {
    const int op = P.insns[pc].code >> 4, imm = P.insns[pc].imm;
    if (op == VM_LOAD) { u.used[imm] = true; u.guarded[imm] = false; }
    if (op == VM_LOADG) u.used[imm] = true;
    if (op == VM_MASK) u.mused[imm] = true;
  }
  int next[kVmMaxLoads], has_prev[kVmMaxLoads] = {};
  for (int s = 0; s < kVmMaxLoads; ++s) next[s] = -1;
  for (int s = 0; s < kVmMaxLoads; ++s) {
    if (!u.used[s] || u.guarded[s]) continue;
    for (int t = 0; t < kVmMaxLoads; ++t) {
      if (t == s || !u.used[t] || u.guarded[t] || P.loads[t].cls != P.loads[s].cls) continue;
      if (P.loads[t].base == P.loads[s].base + P.cls_XY[P.loads[s].cls] && next[s] < 0 && !has_prev[t]) {
        next[s] = t;
        has_prev[t] = 1;
      }
    }
  }
  for (int s = 0; s < kVmMaxLoads; ++s) {
    if (!u.used[s] || u.guarded[s] || has_prev[s]) continue;
    std::vector<int> c;
    for (int t = s; t >= 0 && c.size() < (size_t)kVmMaxLoads; t = next[t]) c.push_back(t);
    u.chains.push_back(c);
  }
  return u;
}

// Distinct reads per thread for V cells of a full z chunk.
int reads_per_thread(const SlotUse& u, int V)
{
  int n = 0;
  for (const auto& c : u.chains) n += (int)c.size() + V - 1;
  for (int s = 0; s < kVmMaxLoads; ++s)
    if (u.used[s] && u.guarded[s]) n += V;
  for (int m = 0; m < kVmMaxMasks; ++m)
    if (u.mused[m]) n += V;
  return n;
}

// The expression one stack-program operation forms from the names of its
// operands (in push order) and of its constant.
std::string op_expr(int op, const std::string& k, const std::vector<std::string>& v)
{
  if (op == VM_CONST) return k;
  if (op == VM_MULC) return v[0] + " * " + k;
  if (op == VM_SELECT) return "(" + v[0] + " != 0.0) ? " + v[1] + " : " + v[2];
  if (op >= VM_ADD && op <= VM_OR) {
    const std::string& a = v[0];
    const std::string& b = v[1];
    switch (op) {
    case VM_ADD: return a + " + " + b;
    case VM_SUB: return a + " - " + b;
    case VM_MUL: return a + " * " + b;
    case VM_DIV: return a + " / " + b;
    case VM_MAX: return "(" + a + " > " + b + ") ? " + a + " : " + b;
    case VM_MIN: return "(" + a + " < " + b + ") ? " + a + " : " + b;
    case VM_POW: return "pow(" + a + ", " + b + ")";
    case VM_EQ: return "(" + a + " == " + b + ") ? 1.0 : 0.0";
    case VM_NE: return "(" + a + " != " + b + ") ? 1.0 : 0.0";
    case VM_LT: return "(" + a + " < " + b + ") ? 1.0 : 0.0";
    case VM_LE: return "(" + a + " <= " + b + ") ? 1.0 : 0.0";
    case VM_GT: return "(" + a + " > " + b + ") ? 1.0 : 0.0";
    case VM_GE: return "(" + a + " >= " + b + ") ? 1.0 : 0.0";
    case VM_AND: return "(" + a + " != 0.0 && " + b + " != 0.0) ? 1.0 : 0.0";
    default: return "(" + a + " != 0.0 || " + b + " != 0.0) ? 1.0 : 0.0";   // VM_OR
    }
  }
  const std::string& a = v[0];
  switch (op) {
  case VM_NEG: return "-" + a;
  case VM_ABS: return "fabs(" + a + ")";
  case VM_SQRT: return "__builtin_sqrt(" + a + ")";
  case VM_SQUARE: return a + " * " + a;
  case VM_CUBE: return a + " * " + a + " * " + a;
  case VM_POWC: return "pow(" + a + ", " + k + ")";
  case VM_EXP: return "exp(" + a + ")";
  case VM_LOG: return "log(" + a + ")";
  case VM_LOG10: return "log10(" + a + ")";
  case VM_SIN: return "sin(" + a + ")";
  case VM_COS: return "cos(" + a + ")";
  case VM_TAN: return "tan(" + a + ")";
  case VM_TANH: return "tanh(" + a + ")";
  case VM_ERF: return "erf(" + a + ")";
  case VM_ERFC: return "erfc(" + a + ")";
  case VM_INV_ERF: return "erfinv(" + a + ")";
  case VM_INV_ERFC: return "erfcinv(" + a + ")";
  default: return "((" + a + ") != 0.0) ? 0.0 : 1.0";   // VM_NOT
  }
}

// Operands an operation pops.
int op_arity(int op)
{
  if (op == VM_LOAD || op == VM_LOADG || op == VM_CONST || op == VM_MASK) return 0;
  if (op == VM_SELECT) return 3;
  if (op >= VM_ADD && op <= VM_OR) return 2;
  return 1;
}

// Index set-up, reads and operations for V cells of one z column at (x, y,
// z0 .. z{V-1}); the generated code expects x, y and the z variables in scope
// and returns the name of each cell's value.  With full, the V cells are
// consecutive planes (z{v} = z0 + v) and reads along a z chain are issued once
// per distinct plane; otherwise (a chunk cut by the box end, z clamped) every
// cell reads for itself.
void emit_reads(std::ostringstream& o, const VmProgram& P, const SlotUse& u, int V, bool full, const std::string& ind)
{
  const bool* used = u.used;
  const bool* guarded = u.guarded;
  const bool* mused = u.mused;
  for (int v = 0; v < V; ++v)
    for (int k = 0; k < P.n_classes; ++k)
      o << ind << "const int i" << v << "_" << k << " = x + A.clsX[" << k << "] * y + A.clsXY[" << k << "] * z"
        << v << ";\n";
  // phase 1: every read, for every cell
  if (full) {
    for (size_t c = 0; c < u.chains.size(); ++c) {
      const std::vector<int>& ch = u.chains[c];
      const int k = P.loads[ch[0]].cls;
      const int planes = (int)ch.size() + V - 1;
      for (int p = 0; p < planes; ++p)
        o << ind << "const double Z" << c << "_" << p << " = A.base[" << ch[0] << "][i0_" << k << " + " << p
          << " * A.clsXY[" << k << "]];\n";
      for (int v = 0; v < V; ++v)
        for (size_t j = 0; j < ch.size(); ++j)
          o << ind << "const double L" << v << "_" << ch[j] << " = Z" << c << "_" << v + (int)j << ";\n";
    }
  }
  for (int v = 0; v < V; ++v) {
    for (int s = 0; s < kVmMaxLoads; ++s) {
      if (!used[s] || (full && !guarded[s])) continue;
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
}

std::vector<std::string> emit_cells(std::ostringstream& o, const VmProgram& P, const SlotUse& u, int V, bool full,
                                    const std::string& ind)
{
  emit_reads(o, P, u, V, full, ind);
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
    const std::string vs = std::to_string(v);
    for (int pc = 0; pc < P.n_insns; ++pc) {
      const int op = P.insns[pc].code >> 4, imm = P.insns[pc].imm;
      if (op == VM_LOAD || op == VM_LOADG) { st.push_back("L" + vs + "_" + std::to_string(imm)); continue; }
      if (op == VM_MASK) { st.push_back("M" + vs + "_" + std::to_string(imm)); continue; }
      const int n = op_arity(op);
      std::vector<std::string> args(st.end() - n, st.end());
      st.resize(st.size() - n);
      fresh(op_expr(op, "A.c[" + std::to_string(imm) + "]", args));
    }
    // a bare leaf has no temporary of its own; the leaf's name is the value
    results.push_back(st.empty() ? std::string("0.0") : st.back());
  }
  return results;
}

// One z chunk of V cells at zb with nv of them inside the box.  A full chunk
// (nv == V, uniform over the block) reads all V cells at once with shared
// z-chain reads; a chunk cut by the box end evaluates its cells one at a time,
// each in its own scope under v < nv, so the tail never needs more registers
// than one cell.  sink(indent, z variable, value) consumes each cell in z order.
template <class Sink>
void emit_chunk(std::ostringstream& o, const VmProgram& P, const SlotUse& u, int V, const std::string& ind,
                Sink sink)
{
  const std::string in2 = V > 1 ? ind + "  " : ind;
  if (V > 1) o << ind << "if (nv == " << V << ") {\n";
  for (int v = 0; v < V; ++v) o << in2 << "const int z" << v << " = zb + " << v << ";\n";
  std::vector<std::string> r = emit_cells(o, P, u, V, true, in2);
  for (int v = 0; v < V; ++v) sink(in2, "z" + std::to_string(v), r[v]);
  if (V > 1) {
    o << ind << "} else {\n";
    const std::string in3 = in2 + "  ";
    for (int v = 0; v < V - 1; ++v) {   // nv < V: cell V - 1 is never inside
      o << in2 << "if (" << v << " < nv) {\n";
      o << in3 << "const int z0 = zb + " << v << ";\n";
      r = emit_cells(o, P, u, 1, false, in3);
      sink(in3, "z0", r[0]);
      o << in2 << "}\n";
    }
    o << ind << "}\n";
  }
}

// Column schedule for trees whose unguarded reads are stencils: the reads of
// one array (one index class) form a family whose bases differ by
// dx + dy*X + dz*XY with each offset in [-2, 2] and a span of at most 2 per
// axis.  A wave covers 64 consecutive x, each lane RY rows, and walks a z chunk;
// a lane keeps, per family, its own column's values of the rows and planes the
// next z step needs in registers (one new plane loaded per step), x neighbours
// come from the adjacent lanes (DPP), and the values past the wave's two edges
// come from one load per family (and row / plane shift) for all rows.  Reads are
// raw buffer loads bounded by the family's allocation, so lanes and rows past
// the box read 0 past it instead of faulting; their results are not stored.
struct ColumnPlan {
  int nf = 0, RY = 4, PF = 1;   // rows per lane; planes loaded ahead of the last plane in use
  bool pair = false;            // two adjacent columns per lane (16-byte loads and stores)
  bool drop[kVmMaxLoads] = {};  // per family: its lowest plane is read only through z carries
  int root[kVmMaxLoads], cx[kVmMaxLoads], ylo[kVmMaxLoads], yhi[kVmMaxLoads], zlo[kVmMaxLoads], zhi[kVmMaxLoads];
  int fam[kVmMaxLoads], dx[kVmMaxLoads], dy[kVmMaxLoads], dz[kVmMaxLoads];
};
constexpr int kColumnRegs = 96;    // ring doubles per lane, over all families
constexpr int kColumnPrefetch = 1;  // planes loaded ahead, when the ring allows it
// Two columns per lane where the ring fits 2 or more rows that way, for trees of
// two or more families on boxes wider than 64 columns.  At 256^3 (one process,
// clocks settled): div(k grad phi) 67.5 us with two columns and 4 stacked waves
// against 71.5 with one column and one wave (the hand-written kernel 67.2); the
// one-family Laplacian 43.1 us with one column against 44.6 with two.
constexpr bool kColumnPairs = true;
// The plane below a family's ring, when only z carries read it, may be loaded
// once for the chunk's first plane instead of riding in the ring (select_ring).
constexpr bool kColumnDropPlanes = true;
// Defaults from sweeps at 256^3 (div(k grad phi) and Laplacian trees): with one
// column per lane, one wave per block measured as fast as 2-4 waves and gives
// small boxes the most blocks; two columns per lane run 4 waves stacked in y
// (div(k grad phi) 67.5 us against 69.9 with one).  g_column_waves 0 takes
// these; g_column_zc 0 picks the planes per block from the box size.
constexpr int kDefaultColumnWaves = 1, kDefaultPairWaves = 4, kDefaultColumnZc = 8;
std::atomic<int> g_column{1}, g_column_waves{0}, g_column_zc{0};

long round_div(long a, long b) { return a >= 0 ? (a + b / 2) / b : -((-a + b / 2) / b); }

// The families of a tree's reads and their offsets (the ring's shape is chosen
// by select_ring); false for a tree the column schedule does not take.
bool column_plan(const VmProgram& P, const SlotUse& u, ColumnPlan* cp)
{
  for (int m = 0; m < kVmMaxMasks; ++m)
    if (u.mused[m]) return false;   // mask bits past the box are not addressable
  int off[kVmMaxLoads][3];
  cp->nf = 0;
  for (int s = 0; s < kVmMaxLoads; ++s) {
    cp->fam[s] = -1;
    if (!u.used[s] || u.guarded[s]) continue;
    const int k = P.loads[s].cls;
    const long X = P.cls_X[k], XY = P.cls_XY[k];
    if (X < 1 || XY < X) return false;
    for (int f = 0; f < cp->nf && cp->fam[s] < 0; ++f) {
      const int r = cp->root[f];
      // one family per allocation: the clamp range is the root's allocation
      if (P.loads[r].cls != k || P.loads[r].base + P.loads[r].glo != P.loads[s].base + P.loads[s].glo) continue;
      const long diff = (long)(((intptr_t)P.loads[s].base - (intptr_t)P.loads[r].base) / (intptr_t)sizeof(double));
      const long dz = round_div(diff, XY), rem = diff - dz * XY, dy = round_div(rem, X), dx = rem - dy * X;
      if (dx < -2 || dx > 2 || dy < -2 || dy > 2 || dz < -2 || dz > 2) continue;
      cp->fam[s] = f;
      off[s][0] = (int)dx, off[s][1] = (int)dy, off[s][2] = (int)dz;
    }
    if (cp->fam[s] < 0) {
      cp->root[cp->nf] = s;
      cp->fam[s] = cp->nf++;
      off[s][0] = off[s][1] = off[s][2] = 0;
    }
  }
  if (cp->nf == 0) return false;
  bool stencil = false;
  for (int f = 0; f < cp->nf; ++f) {
    int lo[3] = {2, 2, 2}, hi[3] = {-2, -2, -2};
    for (int s = 0; s < kVmMaxLoads; ++s)
      if (cp->fam[s] == f)
        for (int d = 0; d < 3; ++d) lo[d] = off[s][d] < lo[d] ? off[s][d] : lo[d], hi[d] = off[s][d] > hi[d] ? off[s][d] : hi[d];
    for (int d = 0; d < 3; ++d)
      if (hi[d] - lo[d] > 2) return false;
    cp->cx[f] = lo[0] + (hi[0] - lo[0] == 2 ? 1 : 0);
    cp->ylo[f] = lo[1], cp->yhi[f] = hi[1], cp->zlo[f] = lo[2], cp->zhi[f] = hi[2];
    stencil = stencil || hi[1] > lo[1] || hi[2] > lo[2];
    for (int s = 0; s < kVmMaxLoads; ++s)
      if (cp->fam[s] == f) cp->dx[s] = off[s][0] - cp->cx[f], cp->dy[s] = off[s][1], cp->dz[s] = off[s][2];
  }
  return stencil;
}

// Tree form of a stack program: post-order (children before their parent, the
// root last), each node its operation, immediate and operand nodes in push order.
struct TNode {
  int op, imm, nk, kid[3];
};

std::vector<TNode> program_tree(const VmProgram& P)
{
  std::vector<TNode> t;
  std::vector<int> st;
  for (int pc = 0; pc < P.n_insns; ++pc) {
    TNode n{P.insns[pc].code >> 4, P.insns[pc].imm, 0, {-1, -1, -1}};
    n.nk = op_arity(n.op);
    for (int i = n.nk - 1; i >= 0; --i) {
      n.kid[i] = st.back();
      st.pop_back();
    }
    t.push_back(n);
    st.push_back((int)t.size() - 1);
  }
  return t;
}

// Subtrees the column schedule evaluates once and reuses: a subtree A whose
// every read is another subtree B's read moved one row down (or one plane down),
// with the same operations and constants, has at row r (plane z) the value B had
// at row r - 1 (plane z - 1), bit for bit: the same operations on the same
// loaded values.  Stencil chains are full of such pairs (the low and high faces
// of a divergence of a face flux), so a lane forms each y face and z face once:
// y faces carried from row to row, z faces from plane to plane.  A is computed
// in the first row (first plane of the chunk) only.  x-shifted pairs are not
// reused: exchanging a value between lanes costs as many instructions as a face.
struct ShiftPlan {
  std::vector<int> from;    // node -> the node B it is taken from, or -1
  std::vector<int> dir;     // 1: B one row down (y), 2: B one plane down (z), 3: B the lane's first column (x)
  std::vector<int> carry;   // node B -> its carry index, or -1
  std::vector<int> cdir;    // carry index -> 1 (row to row), 2 (plane to plane), 3 (column to column)
};

ShiftPlan shift_plan(const VmProgram& P, const ColumnPlan& cp, const std::vector<TNode>& t)
{
  const int n = (int)t.size();
  ShiftPlan sp;
  sp.from.assign(n, -1);
  sp.dir.assign(n, 0);
  sp.carry.assign(n, -1);
  std::vector<int> first(n), lo(3 * n, 0);
  std::vector<char> ok(n, 0), loads(n, 0);
  for (int i = 0; i < n; ++i) {
    const TNode& a = t[i];
    first[i] = a.nk ? first[a.kid[0]] : i;
    if (a.op == VM_LOAD) {
      const int s = a.imm;
      ok[i] = cp.fam[s] >= 0;
      loads[i] = 1;
      lo[3 * i] = cp.dx[s], lo[3 * i + 1] = cp.dy[s], lo[3 * i + 2] = cp.dz[s];
      continue;
    }
    if (a.op == VM_LOADG || a.op == VM_MASK) continue;   // per-cell reads: never moved
    ok[i] = 1;
    for (int k = 0; k < a.nk; ++k) {
      const int c = a.kid[k];
      ok[i] = ok[i] && ok[c];
      if (!loads[c]) continue;
      for (int d = 0; d < 3; ++d)
        lo[3 * i + d] = loads[i] ? std::min(lo[3 * i + d], lo[3 * c + d]) : lo[3 * c + d];
      loads[i] = 1;
    }
  }
  // signature of a subtree with its reads relative to the subtree's lowest corner
  std::function<void(int, const int*, std::string&)> ser = [&](int i, const int* c, std::string& o) {
    const TNode& a = t[i];
    if (a.op == VM_LOAD) {
      const int s = a.imm;
      o += "L" + std::to_string(cp.fam[s]) + ":" + std::to_string(cp.dx[s] - c[0]) + "," +
           std::to_string(cp.dy[s] - c[1]) + "," + std::to_string(cp.dz[s] - c[2]) + ";";
      return;
    }
    o += "(" + std::to_string(a.op) + "," + std::to_string(a.imm);
    for (int k = 0; k < a.nk; ++k) {
      o += " ";
      ser(a.kid[k], c, o);
    }
    o += ")";
  };
  std::vector<std::string> sig(n);
  for (int i = 0; i < n; ++i)
    if (ok[i] && loads[i] && t[i].nk) ser(i, &lo[3 * i], sig[i]);
  std::vector<char> claimed(n, 0);
  auto claim = [&](int i) {
    for (int j = first[i]; j <= i; ++j) claimed[j] = 1;
  };
  for (int b = n - 1; b >= 0; --b) {   // parents before their children: largest subtrees first
    if (claimed[b] || sig[b].empty()) continue;
    for (int a = n - 1; a >= 0; --a) {
      if (a == b || claimed[a] || sig[a] != sig[b]) continue;
      if ((a >= first[b] && a <= b) || (b >= first[a] && b <= a)) continue;
      const int dx = lo[3 * b] - lo[3 * a], dy = lo[3 * b + 1] - lo[3 * a + 1], dz = lo[3 * b + 2] - lo[3 * a + 2];
      const int d = dx == 0 && dy == 1 && dz == 0   ? 1
                    : dx == 0 && dy == 0 && dz == 1 ? 2
                    : dx == 1 && dy == 0 && dz == 0 && cp.pair ? 3 : 0;
      if (!d) continue;
      sp.from[a] = b;
      sp.dir[a] = d;
      sp.carry[b] = (int)sp.cdir.size();
      sp.cdir.push_back(d);
      claim(a);
      claim(b);
      break;
    }
  }
  return sp;
}

// Families whose lowest plane is read only inside subtrees that take their value
// from the plane below (z carries): the chunk's first plane forms those values
// before the z loop from one load of that plane, so the ring holds one plane
// less of them.
void drop_planes(const ColumnPlan& cp, const std::vector<TNode>& t, const ShiftPlan& sp, bool* drop)
{
  const int n = (int)t.size();
  std::vector<int> first(n);
  std::vector<char> carried(n, 0);
  for (int i = 0; i < n; ++i) first[i] = t[i].nk ? first[t[i].kid[0]] : i;
  for (int a = 0; a < n; ++a)
    if (sp.from[a] >= 0 && sp.dir[a] == 2)
      for (int j = first[a]; j <= a; ++j) carried[j] = 1;
  for (int f = 0; f < cp.nf; ++f) drop[f] = kColumnDropPlanes && cp.zhi[f] > cp.zlo[f];
  for (int i = 0; i < n; ++i) {
    if (t[i].op != VM_LOAD) continue;
    const int f = cp.fam[t[i].imm];
    if (f >= 0 && cp.dz[t[i].imm] == cp.zlo[f] && !carried[i]) drop[f] = false;
  }
}

// The ring's shape: rows per lane, planes ahead, one or two columns per lane
// (pairs: allowed), and the planes each family keeps.  Two columns per lane
// where 4 or 2 rows fit, for trees of two or more families, else one column
// with 4, 2 or 1 rows; false when not even one row fits.
bool select_ring(const VmProgram& P, ColumnPlan* cp, bool pairs)
{
  const std::vector<TNode> tree = program_tree(P);
  // ring doubles per lane at ry rows and pf planes ahead: per family, its rows
  // (two values each with two columns per lane, and its edge value when it
  // reads x neighbours) times its planes in use + pf
  auto ring_regs = [&](int ry, int pf, bool pair, const bool* drop) {
    int n = 0;
    for (int f = 0; f < cp->nf; ++f) {
      bool edge = false;
      for (int s = 0; s < kVmMaxLoads; ++s) edge = edge || (cp->fam[s] == f && cp->dx[s] != 0);
      n += ((ry + cp->yhi[f] - cp->ylo[f]) * (pair ? 2 : 1) + (edge ? 1 : 0)) *
           (cp->zhi[f] - cp->zlo[f] + 1 - (drop[f] ? 1 : 0) + pf);
    }
    return n;
  };
  // Preference: two columns per lane, then one; for each, the ring as it is,
  // then with the planes below dropped (which only pays where it lets the
  // layout fit: the 512^3 transport stages 4.31-4.37 ms with two columns per
  // lane and dropped planes against 4.52-4.55 with one column; div(k grad phi)
  // at 256^3 67.5 us with two columns and 2 rows as they are against 72.3-74.0
  // with its planes dropped).
  cp->RY = 0;
  for (int pair = (pairs && cp->nf >= 2) ? 1 : 0; pair >= 0 && !cp->RY; --pair) {
    cp->pair = pair == 1;
    const ShiftPlan sp = shift_plan(P, *cp, tree);   // x carries depend on the layout
    bool drop[kVmMaxLoads] = {};
    drop_planes(*cp, tree, sp, drop);
    const bool none[kVmMaxLoads] = {};
    for (int use = 0; use <= 1 && !cp->RY; ++use) {
      const bool* d = use ? drop : none;
      for (int pf = kColumnPrefetch; pf >= 1 && !cp->RY; --pf)
        for (int ry = 4; ry >= (pair ? 2 : 1) && !cp->RY; ry /= 2)
          if (ring_regs(ry, pf, cp->pair, d) <= kColumnRegs) {
            cp->RY = ry, cp->PF = pf;
            for (int f = 0; f < cp->nf; ++f) cp->drop[f] = d[f];
          }
    }
  }
  if (!cp->RY) cp->pair = false;
  return cp->RY > 0;
}

// Launch shape of the column schedule: waves per block (stacked in y) and z
// planes per block.
struct ColumnShape {
  int waves, zc, xcd;
};

std::string jit_column_source(const VmProgram& P, const SlotUse& u, const ColumnPlan& cp, const ColumnShape& cs)
{
  const int RY = cp.RY, NW = cs.waves;
  SlotUse direct = u;   // guarded reads stay per cell; no z chains
  direct.chains.clear();
  std::ostringstream o;
  o << "struct JitArgs { " NEBO_JIT_STR(NEBO_JIT_ARGS_FIELDS) " };\n";
  // Family reads are raw buffer loads: a resource per plane whose base is the
  // family's lowest read point of that plane (x = lo0 + cx - 1, y = lo1 + ylo)
  // and whose range ends at the family's allocation, plus a per-lane 32-bit byte
  // offset computed once.  A lane or row past the box may form an offset past
  // the allocation; the hardware returns 0 for it instead of faulting, and its
  // results are not stored.
  o << "__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rs_(const double* base, long pb, long ghi)\n{\n"
       "  long n = (ghi - pb) * 8;\n  n = n < 0 ? 0 : n > 0x7ffffff8L ? 0x7ffffff8L : n;\n"
       "  return __builtin_amdgcn_make_buffer_rsrc((void*)(base + pb), (short)0, (int)n, 0x00020000);\n}\n";
  o << "__device__ __forceinline__ double ld_(__amdgpu_buffer_rsrc_t r, unsigned off)\n{\n"
       "  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));\n}\n";
  o << "__device__ __forceinline__ double rdlane_(double v, int j)\n{\n"
       "  const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);\n"
       "  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);\n"
       "  return __hiloint2double(hi, lo);\n}\n";
  const int PW = cp.pair ? 2 : 1;   // columns per lane
  if (cp.pair) {
    // two adjacent doubles in one 16-byte access (8-byte aligned)
    o << "#ifndef NEBO_JIT_D2\ntypedef double nebo_jit_d2_ __attribute__((ext_vector_type(2), aligned(8)));\n"
         "#define NEBO_JIT_D2 nebo_jit_d2_\n#endif\n";
    o << "__device__ __forceinline__ void ld2_(__amdgpu_buffer_rsrc_t r, unsigned off, double& a, double& b)\n{\n"
         "  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);\n"
         "  a = __hiloint2double((int)v[1], (int)v[0]);\n  b = __hiloint2double((int)v[3], (int)v[2]);\n}\n";
    o << "__device__ __forceinline__ void st2_(double* p, double a, double b)\n{\n"
         "  NEBO_JIT_D2 v;\n  v[0] = a;\n  v[1] = b;\n  __builtin_nontemporal_store(v, (NEBO_JIT_D2*)p);\n}\n";
  }
  o << "extern \"C\" __global__ __launch_bounds__(" << 64 * NW << ") void nebo_fused(const JitArgs A)\n{\n";
  if (cs.xcd) {
    o << "  const int L = (int)blockIdx.x, d = L % " << kXcds << ";\n";
    o << "  const int per = A.nblocks / " << kXcds << ", rem = A.nblocks % " << kXcds << ";\n";
    o << "  const int b = d * per + (d < rem ? d : rem) + L / " << kXcds << ";\n";
  } else {
    o << "  const int b = (int)blockIdx.x;\n";
  }
  o << "  const int lane_ = (int)threadIdx.x & 63, w_ = (int)threadIdx.x >> 6;\n";
  o << "  const int ncx = (A.n0 + " << 64 * PW - 1 << ") >> " << (cp.pair ? 7 : 6) << ", ncy = (A.n1 + "
    << RY * NW - 1 << ") / " << RY * NW << ";\n";
  o << "  const int bz = b / (ncx * ncy), bxy = b - bz * (ncx * ncy), by = bxy / ncx, bx = bxy - by * ncx;\n";
  if (cp.pair) o << "  const int x = A.lo0 + bx * 128 + 2 * lane_, xc_ = x;\n";
  else o << "  const int x = A.lo0 + bx * 64 + lane_;\n";
  o << "  const int y0 = A.lo1 + (by * " << NW << " + w_) * " << RY << ";\n";
  o << "  const int zs = A.lo2 + bz * " << cs.zc << ";\n";
  o << "  const int zend = zs + " << cs.zc << " < A.lo2 + A.n2 ? zs + " << cs.zc << " : A.lo2 + A.n2;\n";
  o << "  const bool xin_ = x < A.lo0 + A.n0;\n  const int ye_ = A.lo1 + A.n1;\n";
  if (cp.pair) o << "  const bool xin1_ = x + 1 < A.lo0 + A.n0;\n";
  // ring value of family f, row j, plane slot k (column h of the lane's two)
  auto ring = [&](int f, int j, int k, int h = 0) {
    return "R" + std::to_string(f) + "_" + std::to_string(j) + "_" + std::to_string(k) +
           (cp.pair ? "_" + std::to_string(h) : std::string());
  };
  // Edge reads: per family, one more load per plane serves every x-shifted read
  // of that plane: lanes 0 .. nE-1 read column x0 - 1 of rows ey .. ey+nE-1 (ey
  // the lowest dy of the family's x-shifted reads, nE = RY + their dy span) when
  // the family reads x - 1, lanes 32 .. 32+nE-1 column x0 + 64 (x0 + 128 with
  // two columns per lane) when it reads x + 1; other lanes re-read a ring row.  The value travels with its plane in
  // the register ring, loaded the same number of planes ahead as the rows.  A
  // side the family does not read may lie outside the region its reads cover,
  // below the allocation's start where the range check does not reach (a
  // family reading x and x + 1 only, on a field whose assignment starts at the
  // block's first column), so it is never loaded.
  struct EdgeFam { bool left = false, right = false; int ey = 2, eyhi = -2; };
  std::vector<EdgeFam> ef(cp.nf);
  for (int s = 0; s < kVmMaxLoads; ++s) {
    if (cp.fam[s] < 0 || cp.dx[s] == 0) continue;
    EdgeFam& e = ef[cp.fam[s]];
    (cp.dx[s] < 0 ? e.left : e.right) = true;
    e.ey = std::min(e.ey, cp.dy[s]);
    e.eyhi = std::max(e.eyhi, cp.dy[s]);
  }
  auto has_edge = [&](int f) { return ef[f].left || ef[f].right; };
  // per-lane byte offsets from a plane base: ring rows, and the edge read
  for (int f = 0; f < cp.nf; ++f) {
    const int r = cp.root[f], cls = P.loads[r].cls;
    const int rows = RY + cp.yhi[f] - cp.ylo[f];
    o << "  const long pb" << f << "_ = (long)(A.lo0 + " << cp.cx[f] - 1 << ") + (long)A.clsX[" << cls
      << "] * (A.lo1 + " << cp.ylo[f] << ");\n";
    for (int j = 0; j < rows; ++j)
      o << "  const unsigned o" << f << "_" << j << " = (unsigned)((x - A.lo0 + 1) + A.clsX[" << cls
        << "] * (y0 - A.lo1 + " << j << ")) * 8u;\n";
    if (!has_edge(f)) continue;
    const int nE = RY + ef[f].eyhi - ef[f].ey, dyb = ef[f].ey - cp.ylo[f];
    // x of the wave's first lane is x - PW lane_; lane i holds row y0 + ey + i
    const std::string x0 = cp.pair ? "x - 2 * lane_" : "x - lane_";
    o << "  const unsigned eo" << f << "_ = ";
    if (ef[f].left)
      o << "lane_ < " << nE << " ? (unsigned)((" << x0 << " - A.lo0) + A.clsX[" << cls << "] * (y0 - A.lo1 + lane_ + "
        << dyb << ")) * 8u : ";
    if (ef[f].right)
      o << "(lane_ >= 32 && lane_ < " << 32 + nE << ") ? (unsigned)((" << x0 << " + " << 64 * PW + 1
        << " - A.lo0) + A.clsX[" << cls << "] * (y0 - A.lo1 + lane_ - 32 + " << dyb << ")) * 8u : ";
    o << "o" << f << "_" << -cp.ylo[f] << ";\n";
  }
  // the resource of family f at plane zexpr
  auto rsrc = [&](int f, const std::string& zexpr) {
    const int r = cp.root[f], cls = P.loads[r].cls;
    return "plane_rs_(A.base[" + std::to_string(r) + "], pb" + std::to_string(f) + "_ + (long)A.clsXY[" +
           std::to_string(cls) + "] * (" + zexpr + "), (long)A.ghi[" + std::to_string(r) + "])";
  };
  auto edge_reg = [](int f, int k) { return "V" + std::to_string(f) + "_" + std::to_string(k); };
  // ring rows of family f (and its edge read) into plane slot k, plane zexpr
  auto load_plane = [&](const std::string& ind, int f, int k, const std::string& zexpr) {
    const int rows = RY + cp.yhi[f] - cp.ylo[f];
    o << ind << "{\n" << ind << "  const __amdgpu_buffer_rsrc_t rs_ = " << rsrc(f, zexpr) << ";\n";
    for (int j = 0; j < rows; ++j) {
      if (cp.pair)
        o << ind << "  ld2_(rs_, o" << f << "_" << j << ", " << ring(f, j, k, 0) << ", " << ring(f, j, k, 1) << ");\n";
      else
        o << ind << "  " << ring(f, j, k) << " = ld_(rs_, o" << f << "_" << j << ");\n";
    }
    if (has_edge(f)) o << ind << "  " << edge_reg(f, k) << " = ld_(rs_, eo" << f << "_);\n";
    o << ind << "}\n";
  };
  // zl(f): the lowest plane the ring keeps (zlo, or zlo + 1 when the plane below
  // is read only for the chunk's first plane); nslots(f) = np + PF register
  // slots per family (np planes z + zl .. z + zhi in use, the next PF - 1 the
  // planes after them, the last receiving plane z + zl + nslots - 1 while step z
  // computes; loads past the chunk end are not used; past the allocation they
  // return 0)
  auto zl = [&](int f) { return cp.zlo[f] + (cp.drop[f] ? 1 : 0); };
  auto nslots = [&](int f) { return cp.zhi[f] - zl(f) + 1 + cp.PF; };
  for (int f = 0; f < cp.nf; ++f) {
    const int rows = RY + cp.yhi[f] - cp.ylo[f], ns = nslots(f);
    for (int j = 0; j < rows; ++j)
      for (int k = 0; k < ns; ++k)
        for (int h = 0; h < PW; ++h) o << "  double " << ring(f, j, k, h) << " = 0.0;\n";
    if (has_edge(f))
      for (int k = 0; k < ns; ++k) o << "  double " << edge_reg(f, k) << " = 0.0;\n";
    for (int k = 0; k + 1 < ns; ++k) load_plane("  ", f, k, "zs + " + std::to_string(zl(f) + k));
  }
  // Cells: the program as a tree, with the subtrees of shift_plan taken from the
  // row above or the plane below instead of recomputed.  Carries are written
  // after the whole cell, so a subtree taken from the previous plane always reads
  // the previous plane's value, whichever operand order the tree has.
  const std::vector<TNode> tree = program_tree(P);
  const ShiftPlan sp = shift_plan(P, cp, tree);
  // carries of cell (row r, column h): per column with two columns per lane
  auto cy_name = [&](int c, int r, int h = 0) {
    return "cy" + std::to_string(c) + "_" + std::to_string(r) + (cp.pair ? "_" + std::to_string(h) : std::string());
  };
  auto cz_name = [&](int c, int r, int h = 0) {
    return "cz" + std::to_string(c) + "_" + std::to_string(r) + (cp.pair ? "_" + std::to_string(h) : std::string());
  };
  auto cx_name = [](int c, int r) { return "cx" + std::to_string(c) + "_" + std::to_string(r); };
  // cell (row r, column h) of the tree, or (root >= 0) that subtree alone
  auto emit_row = [&](const std::string& ind, int r, int h, int root = -1) {
    int tmp = 0;
    std::vector<std::pair<std::string, std::string>> writes;
    std::function<std::string(int, const std::string&, bool)> ev = [&](int i, const std::string& in,
                                                                        bool computed) -> std::string {
      const TNode& a = tree[i];
      if (!computed && sp.from[i] >= 0) {
        const int c = sp.carry[sp.from[i]];
        if (sp.dir[i] == 1 && r > 0) return cy_name(c, r - 1, h);
        if (sp.dir[i] == 3 && h == 1) return cx_name(c, r);   // the lane's first column formed it
        if (sp.dir[i] == 2) return cz_name(c, r, h);   // the first plane's: formed before the z loop
      }
      std::string name;
      if (a.op == VM_LOAD || a.op == VM_LOADG) {
        name = "L0_" + std::to_string(a.imm);
      } else {
        std::vector<std::string> args;
        for (int k = 0; k < a.nk; ++k) args.push_back(ev(a.kid[k], in, false));
        name = "t0_" + std::to_string(tmp++);
        o << in << "const double " << name << " = " << op_expr(a.op, "A.c[" + std::to_string(a.imm) + "]", args)
          << ";\n";
      }
      if (sp.carry[i] >= 0) {
        const int c = sp.carry[i];
        if (sp.cdir[c] == 2) writes.emplace_back(cz_name(c, r, h), name);
        else if (sp.cdir[c] == 3) {
          if (h == 0) writes.emplace_back(cx_name(c, r), name);
        } else if (r + 1 < RY) {
          writes.emplace_back(cy_name(c, r, h), name);
        }
      }
      return name;
    };
    const std::string res = root >= 0 ? ev(root, ind, true) : ev((int)tree.size() - 1, ind, false);
    for (const auto& w : writes) o << ind << w.first << " = " << w.second << ";\n";
    return res;
  };
  // the plane below's values of a family whose ring starts one plane higher,
  // loaded once for the chunk's first plane
  auto low_reg = [&](int f, int j, int h) {
    return "P" + std::to_string(f) + "_" + std::to_string(j) + (cp.pair ? "_" + std::to_string(h) : std::string());
  };
  // The reads of cell (row r, column h) as L0_<slot>, at step u (ring slots
  // renamed by u) and plane zq + zshift: -1 for the chunk's first plane's carried
  // subtrees, formed at plane zs - 1 before the loop.  A read of a plane the
  // ring does not keep is bound only there (from the plane-below registers).
  auto bind_reads = [&](const std::string& ind, int r, int h, int u, int zshift) {
    auto slot = [&](int f, int k) { return (k + u) % nslots(f); };
    for (int s = 0; s < kVmMaxLoads; ++s) {
      const int f = cp.fam[s];
      if (f < 0) continue;
      const int dz = cp.dz[s] + zshift;
      if (dz < zl(f) && !(zshift < 0 && cp.drop[f] && dz == cp.zlo[f])) continue;
      const bool low = dz < zl(f);
      const int j = r + cp.dy[s] - cp.ylo[f], k = low ? 0 : slot(f, dz - zl(f));
      auto val = [&](int hh) { return low ? low_reg(f, j, hh) : ring(f, j, k, hh); };
      const std::string edge = low ? "PV" + std::to_string(f) : edge_reg(f, k);
      const int e = (cp.dx[s] > 0 ? 32 : 0) + r + cp.dy[s] - ef[f].ey;   // lane holding the edge value
      if (cp.dx[s] == 0) {
        o << ind << "const double L0_" << s << " = " << val(h) << ";\n";
        continue;
      }
      if (cp.pair && (cp.dx[s] < 0) == (h == 1)) {   // the neighbour is the lane's other column
        o << ind << "const double L0_" << s << " = " << val(1 - h) << ";\n";
        continue;
      }
      // the exchange and the edge lane's read are taken by every lane, outside
      // any branch: a lane read inside the select would run with only the edge
      // lane active, and a value the compiler keeps elsewhere (a spill) would be
      // brought back for that lane alone, not for the lane read
      o << ind << "const double D0_" << s << " = __builtin_amdgcn_mov_dpp(" << val(cp.dx[s] < 0 ? PW - 1 : 0)
        << ", " << (cp.dx[s] > 0 ? "0x130" : "0x138") << ", 0xf, 0xf, true);\n";
      o << ind << "const double E0_" << s << " = rdlane_(" << edge << ", " << e << ");\n";
      o << ind << "const double L0_" << s << " = lane_ == " << (cp.dx[s] > 0 ? 63 : 0) << " ? E0_" << s << " : D0_"
        << s << ";\n";
    }
  };
  // One z step at plane zq: the loads of the plane PF planes ahead, then RY
  // cells.  Ring slot k of family f (plane z + zlo + k; the last slot is the
  // plane loading) lives in register (k + u) mod nslots(f), so the unrolled
  // steps u = 0 .. U-1 rename registers instead of copying them; the remainder
  // loop (u = 0) copies.
  auto step = [&](const std::string& in, int u) {
    auto slot = [&](int f, int k) { return (k + u) % nslots(f); };
    for (int f = 0; f < cp.nf; ++f)
      load_plane(in, f, slot(f, nslots(f) - 1), "zq + " + std::to_string(zl(f) + nslots(f) - 1));
    for (size_t c = 0; c < sp.cdir.size(); ++c)
      if (sp.cdir[c] == 1)
        for (int r = 0; r + 1 < RY; ++r)
          for (int h = 0; h < PW; ++h) o << in << "double " << cy_name((int)c, r, h) << ";\n";
    for (int r = 0; r < RY; ++r) {
      const std::string ind = in + "  ";
      o << in << "{\n" << ind << "const int y = y0 + " << r << ";\n" << ind << "const int z0 = zq;\n";
      if (!cp.pair) {
        bind_reads(ind, r, 0, u, 0);
        emit_reads(o, P, direct, 1, true, ind);
        const std::string res = emit_row(ind, r, 0);
        o << ind << "if (xin_ && y < ye_) __builtin_nontemporal_store(" << res
          << ", &A.out[x + A.oX * y + A.oXY * zq]);\n" << in << "}\n";
        continue;
      }
      // two columns: the first forms the x faces the second takes over
      for (size_t c = 0; c < sp.cdir.size(); ++c)
        if (sp.cdir[c] == 3) o << ind << "double " << cx_name((int)c, r) << ";\n";
      o << ind << "double v0_, v1_;\n";
      for (int h = 0; h < 2; ++h) {
        const std::string in2 = ind + "  ";
        o << ind << "{\n" << in2 << "const int x = xc_ + " << h << ";\n";
        bind_reads(in2, r, h, u, 0);
        emit_reads(o, P, direct, 1, true, in2);
        const std::string res = emit_row(in2, r, h);
        o << in2 << "v" << h << "_ = " << res << ";\n" << ind << "}\n";
      }
      o << ind << "if (y < ye_) {\n"
        << ind << "  if (xin1_) st2_(&A.out[x + A.oX * y + A.oXY * zq], v0_, v1_);\n"
        << ind << "  else if (xin_) __builtin_nontemporal_store(v0_, &A.out[x + A.oX * y + A.oXY * zq]);\n"
        << ind << "}\n" << in << "}\n";
    }
  };
  int U = 1;
  for (int f = 0; f < cp.nf; ++f) {
    const int n = nslots(f);
    int a = U, b = n;
    while (b) { const int t = a % b; a = b; b = t; }
    U = U / a * n;
  }
  if (U > 6) U = 1;   // families of different depths: copies rather than a long unroll
  for (size_t c = 0; c < sp.cdir.size(); ++c)
    if (sp.cdir[c] == 2)
      for (int r = 0; r < RY; ++r)
        for (int h = 0; h < PW; ++h) o << "  double " << cz_name((int)c, r, h) << " = 0.0;\n";
  // The chunk's first plane: its subtrees taken from the plane below are formed
  // at plane zs - 1, from the ring and from one load of each dropped plane.
  bool zcarry = false;
  for (size_t c = 0; c < sp.cdir.size(); ++c) zcarry = zcarry || sp.cdir[c] == 2;
  if (zcarry) {
    o << "  {\n";
    for (int f = 0; f < cp.nf; ++f) {
      if (!cp.drop[f]) continue;
      const int rows = RY + cp.yhi[f] - cp.ylo[f];
      for (int j = 0; j < rows; ++j)
        for (int h = 0; h < PW; ++h) o << "    double " << low_reg(f, j, h) << ";\n";
      if (has_edge(f)) o << "    double PV" << f << ";\n";
      o << "    {\n      const __amdgpu_buffer_rsrc_t rs_ = " << rsrc(f, "zs + " + std::to_string(cp.zlo[f])) << ";\n";
      for (int j = 0; j < rows; ++j) {
        if (cp.pair)
          o << "      ld2_(rs_, o" << f << "_" << j << ", " << low_reg(f, j, 0) << ", " << low_reg(f, j, 1) << ");\n";
        else
          o << "      " << low_reg(f, j, 0) << " = ld_(rs_, o" << f << "_" << j << ");\n";
      }
      if (has_edge(f)) o << "      PV" << f << " = ld_(rs_, eo" << f << "_);\n";
      o << "    }\n";
    }
    for (int r = 0; r < RY; ++r)
      for (int h = 0; h < PW; ++h) {
        o << "    {\n      const int y = y0 + " << r << ";\n      const int z0 = zs - 1;\n";
        if (cp.pair) o << "      const int x = xc_ + " << h << ";\n";
        bind_reads("      ", r, h, 0, -1);
        for (int i = 0; i < (int)tree.size(); ++i)
          if (sp.carry[i] >= 0 && sp.cdir[sp.carry[i]] == 2) {   // each subtree in a scope of its own
            o << "      {\n";
            emit_row("        ", r, h, i);
            o << "      }\n";
          }
        o << "    }\n";
      }
    o << "  }\n";
  }
  o << "  int z = zs;   // column schedule\n";
  if (U > 1) {
    o << "  for (; z + " << U - 1 << " < zend; z += " << U << ") {\n";
    for (int u = 0; u < U; ++u) {
      o << "    {\n      const int zq = z + " << u << ";\n";
      step("      ", u);
      o << "    }\n";
    }
    o << "  }\n";
  }
  o << "  for (; z < zend; ++z) {\n    const int zq = z;\n";
  step("    ", 0);
  for (int f = 0; f < cp.nf; ++f) {
    const int rows = RY + cp.yhi[f] - cp.ylo[f];
    for (int k = 0; k + 1 < nslots(f); ++k) {
      for (int j = 0; j < rows; ++j)
        for (int h = 0; h < PW; ++h) o << "    " << ring(f, j, k, h) << " = " << ring(f, j, k + 1, h) << ";\n";
      if (has_edge(f)) o << "    " << edge_reg(f, k) << " = " << edge_reg(f, k + 1) << ";\n";
    }
  }
  o << "  }\n}\n";
  return o.str();
}

// Assignment kernel: one thread per (x, y) point of the box and V z cells.
// Blocks form a 1-D grid of A.ntx plane tiles times A.nzc z chunks (A.nblocks
// in all), plane tiles fastest.
std::string jit_source(const VmProgram& P, int V, const Shape& sh)
{
  const SlotUse u = slot_use(P);
  std::ostringstream o;
  o << "struct JitArgs { " NEBO_JIT_STR(NEBO_JIT_ARGS_FIELDS) " };\n";
  o << "extern \"C\" __global__ __launch_bounds__(" << sh.block << ") void nebo_fused(const JitArgs A)\n{\n";
  if (sh.xcd) {
    // hardware block L runs on XCD L % 8; give XCD d the contiguous range of
    // logical blocks [start_d, start_d + count_d), count_d = per + (d < rem)
    o << "  const int L = (int)blockIdx.x, d = L % " << kXcds << ";\n";
    o << "  const int per = A.nblocks / " << kXcds << ", rem = A.nblocks % " << kXcds << ";\n";
    o << "  const int b = d * per + (d < rem ? d : rem) + L / " << kXcds << ";\n";
  } else {
    o << "  const int b = (int)blockIdx.x;\n";
  }
  o << "  const int bz = b / A.ntx, bt = b - bz * A.ntx;\n";
  if (sh.tile_x == 0) {
    o << "  const int t = bt * " << sh.block << " + (int)threadIdx.x;\n";
    o << "  if (t >= A.nplane) return;\n";
    o << "  const int yy = t / A.n0;\n  const int x = A.lo0 + (t - yy * A.n0);\n  const int y = A.lo1 + yy;\n";
  } else {
    const int TX = sh.tile_x, TY = sh.block / sh.tile_x;
    o << "  const int tpr = (A.n0 + " << TX - 1 << ") / " << TX << ";\n";
    o << "  const int by = bt / tpr, bx = bt - by * tpr;\n";
    o << "  const int xi = bx * " << TX << " + (int)threadIdx.x % " << TX << ";\n";
    o << "  const int yi = by * " << TY << " + (int)threadIdx.x / " << TX << ";\n";
    o << "  if (xi >= A.n0 || yi >= A.n1) return;\n";
    o << "  const int x = A.lo0 + xi, y = A.lo1 + yi;\n";
  }
  o << "  const int zb = A.lo2 + bz * " << V << ";\n";
  o << "  const int left = A.lo2 + A.n2 - zb;\n  const int nv = left < " << V << " ? left : " << V << ";\n";
  emit_chunk(o, P, u, V, "  ", [&](const std::string& ind, const std::string& z, const std::string& val) {
    o << ind << "__builtin_nontemporal_store(" << val << ", &A.out[x + A.oX * y + A.oXY * " << z << "]);\n";
  });
  o << "}\n";
  return o.str();
}

// Reduction kernel: a grid-stride loop over (x, y, z chunk) items, V z cells
// per item, each thread folding its cells in z order into a double-double
// (sum, norm) or a select-form max/min exactly as the interpreter's
// vm_reduce_kernel does; then a fixed-order block reduction writes one
// (hi, lo) partial per block to A.out, which reduce_final_kernel folds.
std::string jit_reduce_source(const VmProgram& P, int V, int op)
{
  const SlotUse u = slot_use(P);
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
  const int B = kReduceBlock;
  o << "extern \"C\" __global__ __launch_bounds__(" << B << ") void nebo_reduce(const JitArgs A)\n{\n";
  o << "  double hi = " << init << ", lo = 0.0;\n";
  o << "  const long items = (long)A.nplane * A.nzc;\n";
  o << "  for (long w = (long)blockIdx.x * " << B << " + threadIdx.x; w < items; w += (long)A.nblocks * " << B
    << ") {\n";
  o << "    const int zc = (int)(w / A.nplane);\n    const int t = (int)(w - (long)zc * A.nplane);\n";
  o << "    const int yy = t / A.n0;\n    const int x = A.lo0 + (t - yy * A.n0);\n    const int y = A.lo1 + yy;\n";
  o << "    const int zb = A.lo2 + zc * " << V << ";\n";
  o << "    const int left = A.lo2 + A.n2 - zb;\n    const int nv = left < " << V << " ? left : " << V << ";\n";
  emit_chunk(o, P, u, V, "    ", [&](const std::string& ind, const std::string&, const std::string& val) {
    o << ind << "{\n";
    o << ind << "  const double xv = " << val << ";\n";
    if (op == NEBO_RED_MAX) o << ind << "  hi = (hi < xv) ? xv : hi;\n";
    else if (op == NEBO_RED_MIN) o << ind << "  hi = (hi > xv) ? xv : hi;\n";
    else if (op == NEBO_RED_SUM)
      o << ind << "  const double s = hi + xv; const double bb = s - hi;\n"
        << ind << "  lo = lo + ((hi - (s - bb)) + (xv - bb)); hi = s;\n";
    else
      // x*x exactly as p + err (Dekker split, no FMA)
      o << ind << "  const double p = xv * xv; const double t2 = 134217729.0 * xv;\n"
        << ind << "  const double xh = t2 - (t2 - xv), xl = xv - xh;\n"
        << ind << "  const double err = ((xh * xh - p) + 2.0 * xh * xl) + xl * xl;\n"
        << ind << "  const double s = hi + p; const double bb = s - hi;\n"
        << ind << "  lo = lo + ((hi - (s - bb)) + (p - bb)); hi = s; lo = lo + err;\n";
    o << ind << "}\n";
  });
  o << "  }\n";
  // fixed-order block reduction: xor-shuffle tree inside each wave, then the waves in order
  o << "  for (int off = 32; off >= 1; off >>= 1) {\n";
  o << "    const double h2 = __shfl_xor(hi, off, 64);\n";
  o << "    const double l2 = " << (dd ? "__shfl_xor(lo, off, 64)" : "0.0") << ";\n";
  o << "    " << combine << "\n  }\n";
  o << "  __shared__ double sh[" << 2 * (B / 64) << "];\n";
  o << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n";
  o << "  if (lane == 0) { sh[2 * wave] = hi; sh[2 * wave + 1] = lo; }\n";
  o << "  __syncthreads();\n";
  o << "  if (threadIdx.x == 0) {\n    hi = sh[0]; lo = sh[1];\n";
  o << "    for (int w = 1; w < " << B / 64 << "; ++w) {\n";
  o << "      const double h2 = sh[2 * w], l2 = sh[2 * w + 1];\n      " << combine << "\n    }\n";
  o << "    A.out[2 * blockIdx.x] = hi;\n    A.out[2 * blockIdx.x + 1] = " << (dd ? "lo" : "0.0") << ";\n  }\n";
  o << "}\n";
  return o.str();
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
int get_kernel(const std::string& src, const char* name, Entry* out)
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

// Everything the source generators read from a program (opcodes with their
// immediates, the class and bounds check of each read slot, the z chains, the
// masks read) with the cells per thread, the launch shape and the kernel kind
// (-1 assignment, else the reduction op), as a compact key: equal keys give
// equal sources, so a repeated tree finds its kernel without the source text
// being generated again (tens of microseconds of host time per call).
std::string structural_key(const VmProgram& P, int V, const Shape& sh, int kind)
{
  const SlotUse u = slot_use(P);
  std::string k;
  k.reserve(64 + 4 * P.n_insns);
  auto put = [&k](int v) { k.append(reinterpret_cast<const char*>(&v), sizeof v); };
  put(kind);
  put(V);
  put(sh.block);
  put(sh.tile_x);
  put(sh.xcd);
  put(P.n_classes);
  put(P.n_insns);
  for (int pc = 0; pc < P.n_insns; ++pc) put((int)P.insns[pc].code << 16 | (uint16_t)P.insns[pc].imm);
  for (int s = 0; s < kVmMaxLoads; ++s)
    if (u.used[s]) {
      put(s);
      put(P.loads[s].cls);
      put(u.guarded[s] ? 1 : 0);
    }
  put(-1);
  for (int m = 0; m < kVmMaxMasks; ++m) put(u.mused[m] ? 1 : 0);
  for (const auto& c : u.chains) {
    put(-(int)c.size());
    for (int s2 : c) put(s2);
  }
  return k;
}

// Whether an assignment runs on the column schedule (on by default for every
// tree column_plan accepts) and its shape.  Unless set, the planes per block
// are the most of 128, 64, 32, 16 that still give 4 waves per compute unit
// (longer chunks load fewer halo planes; sweeps measured 64 planes fastest at
// 256^3, div(k grad phi) 75.5-76 us against 79-84 at 32, and 128 at 512^3, the
// transport stages 1.51-1.60 ms against 1.56-1.68 at 64), else 8.
bool column_choice(const VmProgram& P, const nebo_limits_t* lim, ColumnPlan* cp, ColumnShape* cs)
{
  const int mode = g_column.load();
  const bool wide = !lim || lim->hi[0] - lim->lo[0] > 64;
  if (!mode || !column_plan(P, slot_use(P), cp) || !select_ring(P, cp, kColumnPairs && mode == 1 && wide))
    return false;
  cs->waves = g_column_waves.load();
  if (cs->waves == 0) cs->waves = cp->pair ? kDefaultPairWaves : kDefaultColumnWaves;
  cs->zc = g_column_zc.load();
  cs->xcd = g_xcd.load();
  // Indices of the kernel's stores and per-lane offsets are 32-bit: lanes, rows
  // and planes past the box reach at most 64 PW + 2 + X (RY waves + 2) +
  // (PF + 2) XY past a slot's last valid index (PW columns per lane), which must
  // not overflow (fields hold fewer than 2^31 points, so only fields within that
  // margin of the limit keep the flattened schedule).
  for (int s = 0; s < kVmMaxLoads; ++s) {
    if (cp->fam[s] < 0) continue;
    const int k = P.loads[s].cls;
    const long reach = (long)P.loads[s].ghi + (cp->pair ? 130 : 66) + (long)P.cls_X[k] * (cp->RY * cs->waves + 2) +
                       (long)(cp->PF + 2) * P.cls_XY[k];
    if (reach >= 2147483647L) return false;
  }
  if (cs->zc == 0) {
    cs->zc = kDefaultColumnZc;
    int device = 0;
    if (lim && hipGetDevice(&device) == hipSuccess) {
      const int xw = cp->pair ? 128 : 64;   // columns per wave
      const long cols = (long)((lim->hi[0] - lim->lo[0] + xw - 1) / xw) *
                        ((lim->hi[1] - lim->lo[1] + cp->RY * cs->waves - 1) / (cp->RY * cs->waves)) * cs->waves;
      const long want = 4L * device_cu_count(device);
      const int nz = lim->hi[2] - lim->lo[2];
      for (int zc : {128, 64, 32, 16})
        if (cs->zc == kDefaultColumnZc && cols * ((nz + zc - 1) / zc) >= want) cs->zc = zc;
    }
  }
  return true;
}

std::string column_key(const VmProgram& P, const ColumnPlan& cp, const ColumnShape& cs)
{
  std::string k = structural_key(P, 1, Shape{64 * cs.waves, 0, cs.xcd}, -2);
  auto put = [&k](int v) { k.append(reinterpret_cast<const char*>(&v), sizeof v); };
  put(cp.RY);
  put(cp.PF);
  put(cp.pair ? 1 : 0);
  put(cs.zc);
  put(cp.nf);
  for (int f = 0; f < cp.nf; ++f) {
    put(cp.root[f]);
    put(cp.cx[f]);
    put(cp.ylo[f]);
    put(cp.yhi[f]);
    put(cp.zlo[f]);
    put(cp.drop[f] ? 1 : 0);
    put(cp.zhi[f]);
  }
  for (int s = 0; s < kVmMaxLoads; ++s)
    if (cp.fam[s] >= 0) {
      put(s);
      put(cp.fam[s]);
      put(cp.dx[s]);
      put(cp.dy[s]);
      put(cp.dz[s]);
    }
  return k;
}

std::unordered_map<std::string, Entry> g_by_key;   // key: device + structural key

// Kernel for a program: by structural key, else generated (gen()), found in or
// added to the source cache, and remembered under the key.
template <class Gen>
int get_kernel_keyed(const std::string& skey, Gen gen, const char* name, Entry* out, bool compile = true)
{
  int device = 0;
  int rc = check_hip(hipGetDevice(&device), "Nebo Fused Kernel");
  if (rc) return rc;
  const std::string key = std::to_string(device) + "|" + skey;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_by_key.find(key);
    if (it != g_by_key.end()) {
      *out = it->second;
      ++g_hits;
      return NEBO_OK;
    }
  }
  Entry e;
  if (!compile) {   // not compiled yet: the caller runs the interpreter
    *out = Entry{};
    return NEBO_OK;
  }
  if ((rc = get_kernel(gen(), name, &e))) return rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_by_key[key] = e;
  }
  bump_plan_epoch();   // calls recorded before now ran without this kernel
  *out = e;
  return NEBO_OK;
}

// Cells per thread: the largest power of two up to the configured maximum
// whose distinct reads per thread (z-chain reads shared) stay within the load
// budget, so a thread's reads are all in flight without spilling.
Shape current_shape() { return Shape{g_block.load(), g_tile_x.load(), g_xcd.load()}; }

int choose_cells(const VmProgram& P)
{
  const SlotUse u = slot_use(P);
  const int budget = g_load_budget.load();
  int V = 1;
  for (int c = 2; c <= g_max_cells.load(); c *= 2)
    if (reads_per_thread(u, c) <= budget) V = c;
  return V;
}

// Cells per thread of a reduction: the assignment's count, doubled (up to 16)
// while a thread issues fewer than 8 loads per item, the reads stay within the
// load budget and the box still gives every thread of a full grid (max_blocks
// blocks) at least two items.  A thread's loop runs its items one after
// another, so on a large box more cells per item keep more loads in flight; a
// small box keeps more threads instead.  At 258^3 this takes max(phi) from
// 28.7-32.1 to 27.3 us (8 cells); sum(phi*k), 8 loads per item at 4 cells,
// measured 46.5-51.3 us at 4 cells and 47.3-48.0 at 8 and keeps 4.
int reduce_cells(const VmProgram& P, const nebo_limits_t& lim, int max_blocks)
{
  const SlotUse u = slot_use(P);
  const long nplane = (long)(lim.hi[0] - lim.lo[0]) * (lim.hi[1] - lim.lo[1]);
  const long n2 = lim.hi[2] - lim.lo[2];
  int V = choose_cells(P);
  while (V < 16 && reads_per_thread(u, V) < 8 && reads_per_thread(u, 2 * V) <= g_load_budget.load() &&
         nplane * ((n2 + 2 * V - 1) / (2 * V)) >= 2L * max_blocks * kReduceBlock)
    V *= 2;
  return V;
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

int launch(const Entry& e, JitArgs* A, unsigned gx, unsigned gy, unsigned block, void* stream)
{
  size_t bytes = sizeof *A;
  void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, A, HIP_LAUNCH_PARAM_BUFFER_SIZE, &bytes, HIP_LAUNCH_PARAM_END};
  return check_hip(hipModuleLaunchKernel(e.fn, gx, gy, 1, block, 1, 1, 0, (hipStream_t)stream, nullptr, cfg),
                   "Nebo Fused Kernel");
}

// A launch of a compiled kernel for the call memo: each replay counts as a
// cache hit, and a reduction's partials go to the replaying call's buffer.
Launch replayed_launch(const Entry& e, const JitArgs& A, unsigned grid, unsigned block, bool reduction)
{
  return [e, A, grid, block, reduction, first = std::make_shared<std::atomic<bool>>(true)](
             void* stream, const CallPtrs& p) {
    if (!first->exchange(false)) ++g_hits;
    JitArgs a = A;
    if (reduction) a.out = p.partial;
    return launch(e, &a, grid, 1, block, stream);
  };
}

}  // namespace

bool jit_enabled() { return g_mode.load() != 0; }

// Arguments and grid of a column-schedule launch of P over lim into lhs.
int column_args(const nebo_field_t* lhs, const VmProgram& P, const nebo_limits_t& lim, const ColumnPlan& cp,
                const ColumnShape& cs, JitArgs* A, long* total)
{
  fill_args(P, lim, 1, A);
  A->out = frame_origin(lhs);
  A->oX = lhs->glob[0];
  A->oXY = lhs->glob[0] * lhs->glob[1];
  const long xw = cp.pair ? 128 : 64;   // columns per wave
  *total = (((long)A->n0 + xw - 1) / xw) * (((long)A->n1 + cp.RY * cs.waves - 1) / (cp.RY * cs.waves)) *
           (((long)A->n2 + cs.zc - 1) / cs.zc);
  if (*total >= (1L << 31)) return set_error(NEBO_ERR_TOO_LARGE, "Nebo Fused Kernel", "more than 2^31 blocks");
  A->nblocks = (int)*total;
  return NEBO_OK;
}

int jit_assign(const nebo_field_t* lhs, const VmProgram& P, const nebo_limits_t& lim, void* stream, bool* done)
{
  *done = false;
  if (!jit_enabled()) return NEBO_OK;
  const int V = choose_cells(P);
  const Shape sh = current_shape();
  Entry e;
  ColumnPlan cp;
  ColumnShape cs;
  if (column_choice(P, &lim, &cp, &cs)) {
    int rc = get_kernel_keyed(column_key(P, cp, cs), [&] { return jit_column_source(P, slot_use(P), cp, cs); },
                              "nebo_fused", &e);
    if (rc || !e.fn) return rc;
    JitArgs A;
    long total = 0;
    if ((rc = column_args(lhs, P, lim, cp, cs, &A, &total))) return rc;
    BackendScope bs("compiled kernel, column schedule (" + std::to_string(cs.waves) + " waves per block, " +
                    std::to_string(cp.RY) + " rows per lane, " + std::to_string(cs.zc) + " planes per block)");
    if ((rc = emit(stream, CallPtrs{}, replayed_launch(e, A, (unsigned)total, (unsigned)(64 * cs.waves), false))))
      return rc;
    *done = true;
    return NEBO_OK;
  }
  int rc = get_kernel_keyed(structural_key(P, V, sh, -1), [&] { return jit_source(P, V, sh); }, "nebo_fused", &e);
  if (rc || !e.fn) return rc;
  JitArgs A;
  fill_args(P, lim, V, &A);
  A.out = frame_origin(lhs);
  A.oX = lhs->glob[0];
  A.oXY = lhs->glob[0] * lhs->glob[1];
  long ntx;
  if (sh.tile_x == 0) {
    ntx = ((long)A.nplane + sh.block - 1) / sh.block;
  } else {
    const long ty = sh.block / sh.tile_x;
    ntx = (((long)A.n0 + sh.tile_x - 1) / sh.tile_x) * (((long)A.n1 + ty - 1) / ty);
  }
  const long total = ntx * A.nzc;
  if (total >= (1L << 31))
    return set_error(NEBO_ERR_TOO_LARGE, "Nebo Fused Kernel", "more than 2^31 blocks");
  A.ntx = (int)ntx;
  A.nblocks = (int)total;
  BackendScope bs("compiled kernel, " + std::to_string(V) + " z cells per thread (" + std::to_string(sh.block) +
                  "-thread blocks)");
  if ((rc = emit(stream, CallPtrs{}, replayed_launch(e, A, (unsigned)total, (unsigned)sh.block, false)))) return rc;
  *done = true;
  return NEBO_OK;
}

int jit_reduce(const VmProgram& P, const nebo_limits_t& lim, int op, const CallPtrs& ptrs, int max_blocks, bool compile,
               void* stream, int* blocks, bool* done)
{
  *done = false;
  if (!jit_enabled()) return NEBO_OK;
  const int V = reduce_cells(P, lim, max_blocks);
  Entry e;
  int rc = get_kernel_keyed(structural_key(P, V, Shape{0, 0, 0}, op), [&] { return jit_reduce_source(P, V, op); },
                            "nebo_reduce", &e, compile);
  if (rc || !e.fn) return rc;
  JitArgs A;
  fill_args(P, lim, V, &A);
  A.out = ptrs.partial;
  const long items = (long)A.nplane * A.nzc;
  long nb = (items + kReduceBlock - 1) / kReduceBlock;
  if (nb > max_blocks) nb = max_blocks;
  if (nb < 1) nb = 1;
  A.nblocks = (int)nb;
  BackendScope bs("compiled reduction kernel (" + std::to_string(nb) + " blocks)");
  if ((rc = emit(stream, ptrs, replayed_launch(e, A, (unsigned)nb, kReduceBlock, true)))) return rc;
  *blocks = (int)nb;
  *done = true;
  return NEBO_OK;
}

std::string jit_source_for(const VmProgram& P, int reduce_op)
{
  const int V = choose_cells(P);
  ColumnPlan cp;
  ColumnShape cs;
  if (reduce_op < 0 && column_choice(P, nullptr, &cp, &cs)) return jit_column_source(P, slot_use(P), cp, cs);
  return reduce_op < 0 ? jit_source(P, V, current_shape()) : jit_reduce_source(P, V, reduce_op);
}

}  // namespace nebo

extern "C" int nebo_set_fused_jit(int enable)
{
  nebo::clear_error();
  nebo::g_mode.store(enable ? 1 : 0);
  nebo::bump_plan_epoch();
  return NEBO_OK;
}

extern "C" int nebo_fused_jit_column(int32_t mode, int32_t waves, int32_t zchunk)
{
  nebo::clear_error();
  static const char* ctx = "Nebo Fused Kernel";
  const int w = waves;   // 0: 4 with two columns per lane, else 1
  const int z = zchunk;   // 0: from the box size
  if (mode < 0 || mode > 2) return nebo::set_error(NEBO_ERR_INVALID, ctx, "column mode must be 0, 1 or 2");
  if (w < 0 || w > 16) return nebo::set_error(NEBO_ERR_INVALID, ctx, "column waves must be in [1, 16] or 0");
  if (z < 0 || z > 4096) return nebo::set_error(NEBO_ERR_INVALID, ctx, "column z chunk must be in [1, 4096] or 0");
  nebo::g_column.store(mode);
  nebo::g_column_waves.store(w);
  nebo::g_column_zc.store(z);
  nebo::bump_plan_epoch();
  return NEBO_OK;
}

extern "C" int nebo_fused_jit_tune(int32_t block_threads, int32_t tile_x, int32_t max_cells,
                                   int32_t load_budget, int32_t xcd_remap)
{
  nebo::clear_error();
  static const char* ctx = "Nebo Fused Kernel";
  const int b = block_threads ? block_threads : nebo::kDefaultBlock;
  const int tx = tile_x >= 0 ? tile_x : nebo::kDefaultTileX;
  const int m = max_cells ? max_cells : nebo::kDefaultMaxCells;
  const int l = load_budget ? load_budget : nebo::kDefaultLoadBudget;
  const int x = xcd_remap >= 0 ? xcd_remap : nebo::kDefaultXcd;
  if (b < 64 || b > 1024 || b % 64)
    return nebo::set_error(NEBO_ERR_INVALID, ctx, "block_threads: a multiple of 64 in [64, 1024]");
  if (tx != 0 && (tx % 64 || b % tx))
    return nebo::set_error(NEBO_ERR_INVALID, ctx, "tile_x: 0, or a multiple of 64 that divides block_threads");
  if (m < 1 || m > 16 || (m & (m - 1)))
    return nebo::set_error(NEBO_ERR_INVALID, ctx, "max_cells: a power of two in [1, 16]");
  if (l < 1 || l > 256) return nebo::set_error(NEBO_ERR_INVALID, ctx, "load_budget: in [1, 256]");
  if (x != 0 && x != 1) return nebo::set_error(NEBO_ERR_INVALID, ctx, "xcd_remap: 0 or 1");
  nebo::g_block.store(b);
  nebo::g_tile_x.store(tx);
  nebo::g_max_cells.store(m);
  nebo::g_load_budget.store(l);
  nebo::g_xcd.store(x);
  nebo::bump_plan_epoch();
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

extern "C" int nebo_fused_column_launch(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost, char* src,
                                        int32_t src_size, int32_t* src_needed, void* args, int32_t args_size,
                                        int32_t* grid, int32_t* block)
{
  static const char* ctx = "Nebo Fused Kernel";
  nebo::clear_error();
  if (!lhs || !rhs) return nebo::set_error(NEBO_ERR_INVALID, ctx, "null argument");
  nebo_limits_t lim;
  int empty = 0;
  int rc = nebo::assign_limits(lhs, rhs, use_ghost, &lim, &empty);
  if (rc) return rc;
  if (empty) return nebo::set_error(NEBO_ERR_UNSUPPORTED, ctx, "empty region: no launch");
  nebo::VmProgram P;
  int depth = 0;
  if ((rc = nebo::lower_program(rhs, &P, &depth))) return rc;
  nebo::ColumnPlan cp;
  nebo::ColumnShape cs;
  if (!nebo::column_choice(P, &lim, &cp, &cs))
    return nebo::set_error(NEBO_ERR_UNSUPPORTED, ctx, "the tree does not run on the column schedule");
  nebo::JitArgs A;
  long total = 0;
  if ((rc = nebo::column_args(lhs, P, lim, cp, cs, &A, &total))) return rc;
  const std::string s = nebo::jit_column_source(P, nebo::slot_use(P), cp, cs);
  if (src_needed) *src_needed = (int32_t)s.size() + 1;
  if (src && src_size > 0) {
    const size_t n = s.size() < (size_t)src_size - 1 ? s.size() : (size_t)src_size - 1;
    std::memcpy(src, s.data(), n);
    src[n] = '\0';
  }
  if (args) {
    if (args_size < (int32_t)sizeof A) return nebo::set_error(NEBO_ERR_INVALID, ctx, "argument buffer too small");
    std::memcpy(args, &A, sizeof A);
  }
  if (grid) *grid = (int32_t)total;
  if (block) *block = 64 * cs.waves;
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
