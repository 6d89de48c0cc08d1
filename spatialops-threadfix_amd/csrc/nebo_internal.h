// nebo_internal.h -- shared declarations of the engine's host and device code.
#pragma once

#include "../../include/nebo_hip.h"

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace nebo {

// ---------------------------------------------------------------------------
// errors: thread-local message, "Nebo error in <ctx>:\n<msg>" like the reference
// ---------------------------------------------------------------------------
int set_error(int code, const char* ctx, const std::string& msg);
void clear_error();

// Names the kernel that evaluates an assignment or reduction: recorded for
// nebo_last_backend (per thread) and, when reporting is on, printed as
// "Starting Nebo HIP <name>" now and "Finished Nebo HIP <name>" when the scope
// ends (the kernel is then queued).
bool report_backend();
struct BackendScope {
  explicit BackendScope(const std::string& name);
  ~BackendScope();
  BackendScope(const BackendScope&) = delete;
  BackendScope& operator=(const BackendScope&) = delete;
};

// ---------------------------------------------------------------------------
// call memo: an assignment or reduction whose descriptors are byte for byte
// those of an earlier successful call (same kind, flags, current device and
// planning settings) replays that call's launches instead of checking, lowering
// and planning the tree again, so a time loop that repeats its expressions pays
// one key comparison per call on the host before its launches.
// ---------------------------------------------------------------------------
// Pointers of one call that its descriptors do not hold: a reduction's result
// slot and its block partials.
struct CallPtrs {
  double* out = nullptr;
  double* partial = nullptr;
};
// One launch of a recorded call, on the stream given at replay.
using Launch = std::function<int(void* stream, const CallPtrs& p)>;
// Runs `launch` now and, while the calling thread records a call, keeps it for
// replay with the backend name in effect.  Every launch of an assignment or
// reduction goes through here, or its replay would miss it.
int emit(void* stream, const CallPtrs& p, Launch launch);
// Drops every recorded call.  Called by whatever changes what planning reads:
// pinned and tuned tiles, the run-time compiler's settings, a kernel compiled.
void bump_plan_epoch();

// Key of a call: its kind, flags and current device, then the bytes of its
// descriptors (only the points a stencil defines).  expr() is false for a
// descriptor it cannot read safely (negative counts, null tables): such calls
// are not recorded, and planning reports their error.
struct CallKey {
  std::string bytes;
  void begin(char kind, int device, int flag_a, int flag_b);
  void field(const nebo_field_t& f);
  bool expr(const nebo_expr_t* e);
};
struct MemoRecord;
// The call recorded under key for the current settings, or null.
std::shared_ptr<const MemoRecord> memo_find(const CallKey& key);
// Replays a recorded call's launches on stream.
int memo_run(const MemoRecord& r, void* stream, const CallPtrs& p);
// Records the launches of the calling thread from construction to commit(),
// which keeps them under the key when the call succeeded.
class MemoRecorder {
 public:
  explicit MemoRecorder(const CallKey& key);
  ~MemoRecorder();
  void commit(int rc);
  MemoRecorder(const MemoRecorder&) = delete;
  MemoRecorder& operator=(const MemoRecorder&) = delete;
 private:
  std::string key_;
  MemoRecord* rec_;
  MemoRecord* prev_;
};

// ---------------------------------------------------------------------------
// ghost algebra (GhostData.h:54-257): six counts; all >= NEBO_GHOST_MAX = infinite
// ---------------------------------------------------------------------------
struct Ghost {
  int m[3];
  int p[3];
  bool infinite() const {
    for (int d = 0; d < 3; ++d)
      if (m[d] < NEBO_GHOST_MAX || p[d] < NEBO_GHOST_MAX) return false;
    return true;
  }
  static Ghost uniform(int n) { return Ghost{{n, n, n}, {n, n, n}}; }
};

// Ghost counts of a node (ghosts_with_bc when with_bc, else ghosts_without_bc).
int node_ghosts(const nebo_expr_t* e, int node, bool with_bc, Ghost* out);
// Validate node indices, opcodes and table indices of an expression.
int validate_expr(const nebo_expr_t* e, const char* ctx);
// Loop limits of lhs <<= rhs (NeboBasic.h:61-100, NeboLhs.h:67-88).
int assign_limits(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                  nebo_limits_t* lim, int* empty);
// Loop limits of a reduction (NeboReductions.h:54-70).
int reduce_limits(const nebo_expr_t* e, int use_ghost, nebo_limits_t* lim, int* empty);

// ---------------------------------------------------------------------------
// generic fused evaluator: postfix stack program with host-known stack depth
// ---------------------------------------------------------------------------
// Opcodes of the device stack machine.  The device dispatches on
// (op << 4) | depth, so every case reads and writes fixed stack slots.
enum VmOp : uint8_t {
  VM_LOAD = 0,   // push field value          imm = load slot
  VM_CONST,      // push constant             imm = const index
  VM_MASK,       // push mask bit (0/1)       imm = mask slot
  VM_MULC,       // top = top * const[imm]    (stencil coefficient)
  VM_ADD, VM_SUB, VM_MUL, VM_DIV, VM_MAX, VM_MIN, VM_POW,
  VM_EQ, VM_NE, VM_LT, VM_LE, VM_GT, VM_GE, VM_AND, VM_OR,
  VM_NEG, VM_ABS, VM_SQRT, VM_SQUARE, VM_CUBE, VM_POWC,  // POWC: pow(top, const[imm])
  VM_EXP, VM_LOG, VM_LOG10, VM_SIN, VM_COS, VM_TAN, VM_TANH, VM_ERF, VM_ERFC, VM_INV_ERF, VM_INV_ERFC,
  VM_NOT,
  VM_SELECT,     // c = pop, b = pop, a = pop, push (a != 0 ? b : c)
  VM_LOADG,      // VM_LOAD that yields NaN where the read leaves the field's block
  VM_NUM_OPS
};

constexpr int kVmMaxDepth = 10;     // stack slots
constexpr int kVmMaxInsns = 200;
constexpr int kVmMaxConsts = 48;
constexpr int kVmMaxLoads = 40;
constexpr int kVmMaxMasks = 6;
constexpr int kVmMaxClasses = 4;    // distinct (glob x, glob y) shapes

struct VmInsn {
  uint16_t code;   // (op << 4) | depth before the instruction
  int16_t imm;
};

struct VmLoad {
  const double* base;   // values + flat(offset + valid_minus + shift)
  int32_t cls;          // index class (glob shape)
  int32_t glo, ghi;     // cell indices i with base + i inside the glob block: [glo, ghi)
  int32_t pad_;
};

struct VmMask {
  const uint32_t* bits;
  int32_t o[3];         // offset + ghost_minus + shift
  int32_t X, Y;
  int32_t pad_;
};

// Kernel-argument image of one lowered expression.
struct VmProgram {
  int32_t n_insns;
  int32_t n_classes;
  int32_t cls_X[kVmMaxClasses];
  int32_t cls_XY[kVmMaxClasses];
  VmLoad loads[kVmMaxLoads];
  VmMask masks[kVmMaxMasks];
  double consts[kVmMaxConsts];
  VmInsn insns[kVmMaxInsns];
};

// Lower a tree to a stack program.  Also records the max stack depth.
int lower_program(const nebo_expr_t* e, VmProgram* prog, int* max_depth);

// Compute units of a GPU (cached per device; 256 if the query fails).
int device_cu_count(int device);

// Fused kernels compiled at run time (nebo_jit.hip): runs the program as one
// generated kernel and sets *done, or leaves *done false when run-time
// compilation is off or failed for this program (the interpreter then runs it).
bool jit_enabled();
int jit_assign(const nebo_field_t* lhs, const VmProgram& P, const nebo_limits_t& lim, void* stream, bool* done);
// Reduction: writes one (hi, lo) partial per block into ptrs.partial (room for
// max_blocks) and the block count into *blocks; the caller folds the partials
// with reduce_final_kernel.  Sets *done, or leaves it false when run-time
// compilation is off, failed for this program, or (compile false) the program's
// kernel has not been compiled by an earlier call.
int jit_reduce(const VmProgram& P, const nebo_limits_t& lim, int op, const CallPtrs& ptrs, int max_blocks, bool compile,
               void* stream, int* blocks, bool* done);

// One (field, offset) read of an expression; edgeless when an edgeless stencil
// lies between the read and the root.
struct FieldRead {
  int field;
  int s[3];
  bool edgeless;
};
// Every distinct read of the subtree at `node` evaluated at offset (sx, sy, sz).
// Call on trees that lower_program accepts (its size limits bound the walk).
void collect_reads(const nebo_expr_t* e, int node, int sx, int sy, int sz, bool edgeless,
                   std::vector<FieldRead>* out);

// ---------------------------------------------------------------------------
// recognised trees
// ---------------------------------------------------------------------------
enum KernelId { K_GENERIC = 0, K_DIVKGRAD = 1, K_LAPL_EULER = 2 };

struct DivKGradMatch {
  int k_field, phi_field;
  int form;                       // 0: (Dx+Dy)+Dz, 1: ((-Dx)-Dy)-Dz
  nebo_axis_coefs_t c;
};
struct LaplEulerMatch {
  int phi_field;                  // read field (centre term and stencils)
  double s;                       // scalar factor
  nebo_axis_coefs_t c;
};
bool match_divkgrad(const nebo_expr_t* e, DivKGradMatch* m);
bool match_lapl_euler(const nebo_expr_t* e, LaplEulerMatch* m);

// ---------------------------------------------------------------------------
// launch helpers (device code)
// ---------------------------------------------------------------------------
int launch_generic_assign(const nebo_field_t* lhs, const VmProgram& prog, int max_depth,
                          const nebo_limits_t& lim, void* stream);
int launch_divkgrad(const nebo_field_t* r, const nebo_field_t* k, const nebo_field_t* phi,
                    const nebo_axis_coefs_t& c, int form, const nebo_limits_t& lim,
                    void* stream);
int launch_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                      const nebo_axis_coefs_t& c, double s, const nebo_limits_t& lim,
                      void* stream);

// out = phi + dt*div(interp(k)*grad(phi)) over lim (the diffusion kernel's Euler form).
int launch_diffusion_euler(const nebo_field_t* out, const nebo_field_t* k, const nebo_field_t* phi,
                           const nebo_axis_coefs_t& c, double dt, const nebo_limits_t& lim,
                           void* stream);
// Periodic copy-fill of every allocated ghost layer of f.
int fill_ghosts_periodic(const nebo_field_t* f, void* stream);

// Periodic operands (diffusion and transport chains): f shares ref's memory
// window and device, has exactly one ghost layer per face (valid) and fewer than
// 2^31 points.  Sets the error with context ctx.
int check_periodic_operand(const nebo_field_t* f, const nebo_field_t* ref, const char* ctx);
// [0, extent - 2) per axis: the interior of a one-ghost periodic field.
nebo_limits_t interior_of(const nebo_field_t* f);
// f's descriptor with another values pointer (same window).
nebo_field_t with_values(const nebo_field_t* f, void* v);
// glob[0] * glob[1] * glob[2]
size_t field_npts(const nebo_field_t* f);

// Pointer of the first valid-ghost-excluded point of a field (the frame origin).
inline double* frame_origin(const nebo_field_t* f)
{
  long gx = f->offset[0] + f->valid_minus[0];
  long gy = f->offset[1] + f->valid_minus[1];
  long gz = f->offset[2] + f->valid_minus[2];
  return static_cast<double*>(f->values) + gx + (long)f->glob[0] * (gy + (long)f->glob[1] * gz);
}

int check_hip(int hip_status, const char* ctx);

}  // namespace nebo
