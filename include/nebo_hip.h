/*
 * nebo_hip.h -- C-ABI of the MI355X (gfx950) evaluation engine for SpatialOps/Nebo
 * field expressions.
 *
 * Every entry point takes plain pointers, sizes and POD descriptors; none of them
 * owns field memory passed in a descriptor.  Every call returns an int status
 * (NEBO_OK == 0); on failure nebo_last_error() returns a thread-local message that
 * starts with "Nebo error in <context>:" so that a C++ wrapper can rethrow it as
 * std::runtime_error exactly like the reference does
 * (spatialops/NeboLhs.h:107-153, spatialops/NeboBasic.h:65-93).
 *
 * Memory layout (identical to the reference, spatialops/structured/MemoryWindow.h:121-131):
 * a field is one x-fastest block of glob[0]*glob[1]*glob[2] doubles; the window is
 * (offset, extent) inside that block; flat index = x + glob[0]*(y + glob[1]*z).
 *
 * Streams are passed as `void*` holding a hipStream_t (NULL = the null stream).
 * No call synchronises implicitly unless its comment says so.  Calls on different
 * streams may run concurrently from different host threads.
 */
#ifndef NEBO_HIP_H
#define NEBO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* status codes                                                              */
/* ------------------------------------------------------------------------- */
#define NEBO_OK               0
#define NEBO_ERR_INVALID      1   /* malformed descriptor / argument            */
#define NEBO_ERR_GHOST        2   /* not enough valid ghost / extra cells       */
#define NEBO_ERR_DEVICE       3   /* operand not resident on the LHS device     */
#define NEBO_ERR_HIP          4   /* HIP runtime error                          */
#define NEBO_ERR_TOO_LARGE    5   /* expression exceeds the evaluator's limits  */
#define NEBO_ERR_UNSUPPORTED  6   /* operation not provided on this device      */

/* device index semantics of spatialops/structured/MemoryTypes.h:14-19 */
#define NEBO_CPU_INDEX (-1)

/* GHOST_MAX of spatialops/structured/GhostData.h:43 ("infinite" ghost count) */
#define NEBO_GHOST_MAX 9001

/* ------------------------------------------------------------------------- */
/* descriptors                                                               */
/* ------------------------------------------------------------------------- */

/* One SpatialField: SpatialField.h:128-603 + FieldInfo.h + BoundaryCellInfo.h. */
typedef struct nebo_field {
  void*   values;          /* first element of the allocation (glob block)      */
  int32_t glob[3];         /* MemoryWindow::glob_dim                            */
  int32_t offset[3];       /* MemoryWindow::offset (window with ghosts)         */
  int32_t extent[3];       /* MemoryWindow::extent (window with ghosts)         */
  int32_t ghost_minus[3];  /* get_ghost_data()       (allocated ghost layers)   */
  int32_t ghost_plus[3];
  int32_t valid_minus[3];  /* get_valid_ghost_data() (valid ghost layers)       */
  int32_t valid_plus[3];
  int32_t has_bc[3];       /* boundary_info().has_bc()                          */
  int32_t bc_extra[3];     /* boundary_info().has_extra()                       */
  int32_t device;          /* active device: NEBO_CPU_INDEX or a GPU ordinal    */
  int32_t pad_;
} nebo_field_t;

/* One SpatialMask (SpatialMask.h:99-181, BitField.h:291-333).  Bit b of the
 * mask is set when the point of flat index b (over the mask's glob block) is in
 * the mask: word b/32, bit b%32. */
typedef struct nebo_mask {
  const uint32_t* bits;    /* device copy of the bit field (add_consumer)       */
  int32_t glob[3];
  int32_t offset[3];       /* window offset (with ghosts)                       */
  int32_t extent[3];       /* window extent (with ghosts)                       */
  int32_t ghost_minus[3];  /* get_ghost_data()                                  */
  int32_t ghost_plus[3];
  int32_t valid_minus[3];  /* get_valid_ghost_data()                            */
  int32_t valid_plus[3];
  int32_t has_bc[3];
  int32_t bc_extra[3];
} nebo_mask_t;

/* Expression node opcodes.  Leaves, pointwise operators (NeboOperators.h),
 * stencils (NeboStencils.h), cond (NeboCond.h) and masks (NeboMask.h). */
enum nebo_opcode {
  NEBO_SCALAR = 0,     /* value                                  NeboRhs.h:29-158  */
  NEBO_FIELD,          /* idx = field table index                NeboRhs.h:160-328 */
  NEBO_MASK,           /* idx = mask table index (boolean)       NeboMask.h:29-213 */
  NEBO_ADD, NEBO_SUB, NEBO_MUL, NEBO_DIV,          /* a (op) b                     */
  NEBO_NEG,                                         /* -a                           */
  NEBO_MAX, NEBO_MIN,   /* (a > b) ? a : b  /  (a < b) ? a : b  NeboOperators.h:13237,13992 */
  NEBO_ABS, NEBO_SQRT, NEBO_SQUARE, NEBO_CUBE,
  NEBO_POW,             /* std::pow(a, b)                                        */
  NEBO_POWINT,          /* std::pow(a, idx)                     NeboOperators.h:14710 */
  NEBO_EXP, NEBO_LOG, NEBO_LOG10, NEBO_SIN, NEBO_COS, NEBO_TAN, NEBO_TANH,
  NEBO_ERF, NEBO_ERFC,
  NEBO_EQ, NEBO_NE, NEBO_LT, NEBO_LE, NEBO_GT, NEBO_GE,   /* boolean results       */
  NEBO_AND, NEBO_OR, NEBO_NOT,                             /* boolean operands     */
  NEBO_STENCIL,         /* a = argument node, idx = stencil table index          */
  NEBO_COND,            /* test a ? value b : otherwise c       NeboCond.h:408-412 */
  NEBO_MASK_SHIFT,      /* a = NEBO_MASK node, idx = stencil table entry whose one
                           point P is the shift: mask(x + P) (boolean)
                           NeboMaskShift, NeboStencils.h:1048-1213             */
  NEBO_INV_ERF,         /* erf^-1(a)  (boost::math::erf_inv)   NeboOperators.h:6150-6330 */
  NEBO_INV_ERFC,        /* erfc^-1(a) (boost::math::erfc_inv), the InvErfcFcn after it */
  NEBO_NUM_OPCODES
};

typedef struct nebo_node {
  int32_t op;            /* enum nebo_opcode                                    */
  int32_t a, b, c;       /* child node indices; every child precedes its parent */
  int32_t idx;           /* field / mask / stencil table index, or exponent     */
  int32_t pad_;
  double  value;         /* NEBO_SCALAR literal                                 */
} nebo_node_t;

#define NEBO_MAX_STENCIL_POINTS 27

/* Stencil kinds (NeboStencils.h) */
enum nebo_stencil_kind {
  NEBO_STENCIL_WEIGHTED = 0,  /* fold_p arg(x+p)*c_p, left fold in point order  :407-448 */
  NEBO_STENCIL_SUM      = 1,  /* fold_p arg(x+p), left fold in point order      :940-975 */
  NEBO_STENCIL_EDGELESS = 2   /* weighted, but the valid range does not shrink on the
                                 side it reads (NeboEdgelessStencil :526-760); reads
                                 outside the operand's allocated block give NaN      */
};

typedef struct nebo_stencil {
  int32_t kind;                               /* enum nebo_stencil_kind     */
  int32_t npts;
  int32_t pts[NEBO_MAX_STENCIL_POINTS][3];    /* offsets, definition order  */
  double  coefs[NEBO_MAX_STENCIL_POINTS];     /* one per point (WEIGHTED)   */
} nebo_stencil_t;

/* A whole right-hand side: node list in post-order (root = last node). */
typedef struct nebo_expr {
  int32_t n_nodes;    const nebo_node_t*    nodes;
  int32_t n_fields;   const nebo_field_t*   fields;
  int32_t n_stencils; const nebo_stencil_t* stencils;
  int32_t n_masks;    const nebo_mask_t*    masks;
} nebo_expr_t;

/* Loop limits in the LHS frame (relative to the first valid-ghost-excluded
 * point): [lo[d], hi[d]) per axis.  NeboLhs.h:79-84. */
typedef struct nebo_limits { int32_t lo[3]; int32_t hi[3]; } nebo_limits_t;

/* ------------------------------------------------------------------------- */
/* runtime                                                                   */
/* ------------------------------------------------------------------------- */

/* Last error message of the calling thread ("" when none). */
const char* nebo_last_error(void);

/* Backend reporting, the reference's NEBO_REPORT_BACKEND messages
 * (spatialops/NeboLhs.h:167-170, 279-282, 321-331, 407, 475): when on, every
 * assignment, masked assignment and reduction prints "Starting Nebo HIP
 * <kernel>" before its launch and "Finished Nebo HIP <kernel>" once it is queued.
 * The reference switches the messages on at compile time; here the
 * NEBO_REPORT_BACKEND environment variable (any value but "0") or this call
 * switches them at run time. */
int nebo_set_report_backend(int on);

/* The kernel the calling thread's last assignment / reduction ran (e.g.
 * "div(k grad phi) kernel, tile 128x8x64 ..."; "" before any). */
const char* nebo_last_backend(void);

/* Number of visible GPUs (CUDADeviceInterface::get_device_count,
 * spatialops/structured/ExternalAllocators.h:59-106). */
int nebo_device_count(int* count);

/* Select `device` and warm the runtime (CUDADeviceInterface::self()). */
int nebo_init(int device);

/* Device memory through a size-bucketed pool that keeps freed blocks for
 * reuse (replaces Pool<T>::get/put, spatialops/structured/MemoryPool.cpp:102-186). */
int nebo_alloc(int device, size_t bytes, void** out);
int nebo_free(int device, void* ptr);
/* Release every pooled block of `device` back to the runtime. */
int nebo_pool_trim(int device);
/* Bytes currently handed out / high-water mark (MemoryPool.h:49,82-90). */
int nebo_pool_stats(int device, size_t* in_use, size_t* high_water);
/* Page-locked host memory, so that host<->device copies on a stream run
 * asynchronously (CUDADeviceInterface::get_pinned_pointer / releaseHost,
 * ExternalAllocators.h:59-106; the reference's pool pins host blocks when CUDA is
 * on, MemoryPool.cpp:132-133).  Fails with NEBO_ERR_HIP where there is no GPU. */
int nebo_host_alloc(size_t bytes, void** out);
int nebo_host_free(void* ptr);

/* Copies.  kind: 0 host->device, 1 device->host, 2 device->device,
 * 3 peer (dst on dst_dev, src on src_dev).  stream NULL => synchronous
 * (CUDADeviceInterface::memcpy_to/from/peer, CudaMemoryAllocator.cu:143-230). */
int nebo_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream);
int nebo_memcpy_peer(void* dst, int dst_dev, const void* src, int src_dev,
                     size_t bytes, void* stream);
int nebo_memset(void* dst, int value, size_t bytes, void* stream);
int nebo_stream_sync(void* stream);
int nebo_device_sync(void);

/* ------------------------------------------------------------------------- */
/* assignment: the seam NeboField<Initial,F>::assign -> gpu_assign            */
/* ------------------------------------------------------------------------- */

/* Host-only: the loop limits `lhs <<= rhs` (use_ghost=1) or
 * `interior_assign(lhs, rhs)` (use_ghost=0) would cover: calculate_actual_ghost
 * (NeboBasic.h:61-100) + NeboLhs.h:67-88.  Sets *empty when the range has zero
 * volume (no kernel is launched).  Throws the reference's ghost errors as
 * NEBO_ERR_GHOST. */
int nebo_assign_limits(const nebo_field_t* lhs, const nebo_expr_t* rhs,
                       int use_ghost, nebo_limits_t* limits, int* empty);

/* lhs <<= rhs on the LHS's GPU: one fused kernel over the whole tree
 * (replaces gpu_assign + gpu_assign_kernel, NeboLhs.h:29-40,400-478).
 * Recognised trees run on hand-written kernels; every other tree runs on the
 * generic fused evaluator.  Results equal the reference CPU backend bit for bit
 * for + - * / neg max min abs sqrt square cube, comparisons, cond, masks and
 * stencils; pow and the transcendental functions (exp log log10 sin cos tan
 * tanh erf erfc) use the device math library and stay within the ulp budgets
 * of the reference's own tests (test/NeboTest.cpp:497-523); pow(x, 2) is
 * evaluated as the correctly rounded x*x. */
int nebo_assign(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                void* stream);

/* Like nebo_assign but forces the generic fused evaluator (no pattern match). */
int nebo_assign_generic(const nebo_field_t* lhs, const nebo_expr_t* rhs,
                        int use_ghost, void* stream);

/* Repeated calls.  An assignment (nebo_assign, nebo_assign_generic) or a
 * reduction (nebo_reduce, nebo_reduce_async, nebo_reduce_async_noalloc) whose
 * descriptors equal those of an earlier successful call byte for byte (the LHS,
 * and the expression's node, field, stencil and mask tables), with the same
 * flags, the same current device and an entry point that plans alike, replays the
 * launches that call planned: the checks, the lowering and the kernel choice are
 * not repeated, so a time loop that keeps its descriptors pays one key
 * comparison per call on the host besides its launches.  Values are read at
 * launch, so new values in the same fields are seen.  The records are dropped
 * whenever something the planning reads changes: nebo_stencil_set_tile,
 * nebo_stencil_autotune, nebo_set_fused_jit, nebo_fused_jit_tune,
 * nebo_fused_jit_column, and every kernel compiled at run time.  The reference
 * has no counterpart: its expression templates make these choices at compile
 * time. */

/* The generic evaluator and the reductions run a tree as one kernel compiled at
 * run time for that tree's shape (hiprtc, gfx950, the library's flags; cached
 * per device and shape: field pointers, constants and sizes are arguments), or,
 * when run-time compilation is switched off or fails, as the stack-program
 * interpreter.  Assignments give the same bits either way; reductions meet the
 * same bounds (max/min exact, sum and norm within one rounding of the exact
 * result) but may order their double-double partial sums differently.  No
 * reference counterpart (the reference instantiates a kernel per expression
 * type at its own compile time).
 *   nebo_set_fused_jit(0 | 1): interpreter only / compiled kernels (default 1);
 *   nebo_fused_jit_stats: kernels compiled, cache hits, failed compiles;
 *   nebo_fused_source (host only): the kernel source generated for rhs, the
 *   assignment kernel for reduce_op -1, else the NEBO_RED_* reduction kernel
 *   (*needed = bytes with the terminating 0; buf may be NULL). */
int nebo_set_fused_jit(int enable);
/* Shape of the compiled assignment kernels: threads per block (a multiple of 64
 * in [64, 1024]); tile_x 0 lays a block over consecutive points of the
 * flattened x-y plane, else over a tile_x by block_threads/tile_x tile (tile_x a
 * multiple of 64 dividing block_threads); z cells per thread: the largest power
 * of two up to max_cells whose distinct reads per thread (reads one plane apart
 * along z shared between cells) stay within load_budget; xcd_remap 1 renumbers
 * blocks so that neighbouring tiles run on one XCD (one L2).  0 keeps the
 * default of block_threads, max_cells and load_budget, -1 that of tile_x and
 * xcd_remap.  Defaults: 512 threads, flattened plane, 4 cells, 48 reads, XCD
 * renumbering on.  Kernels compiled for another shape stay cached. */
int nebo_fused_jit_tune(int32_t block_threads, int32_t tile_x, int32_t max_cells,
                        int32_t load_budget, int32_t xcd_remap);
/* Column schedule of the compiled assignment kernels.  It applies to a tree
 * without masks whose unguarded reads of each field lie within a 3 x 3 x 3
 * neighbourhood and that reads some y or z neighbour.  Each lane holds one
 * column, or two adjacent columns (16-byte loads and stores; the x face its two
 * cells share is formed once) for a tree reading two or more fields on a box
 * wider than 64 columns where the register ring of 96 doubles fits 2 or more
 * rows that way; a wave covers 64 or 128 consecutive x points, each lane 1, 2
 * or 4 rows, walks zchunk planes and loads each plane of each field once, one
 * plane ahead of its use.  x neighbours come from the adjacent lanes.  Other
 * trees keep the shape of nebo_fused_jit_tune.  mode 1 uses it where it
 * applies (the default), 2 the same with one column per lane, 0 never; waves
 * per block (stacked in y) in [1, 16], 0 (the default) for 4 with two columns
 * per lane and 1 with one; zchunk in [1, 4096], 0 (the default) for the most of
 * 128, 64, 32 or 16 planes that still gives 4 waves per compute unit, else 8.
 * Results are the same bits either way. */
int nebo_fused_jit_column(int32_t mode, int32_t waves, int32_t zchunk);
int nebo_fused_jit_stats(int32_t* compiled, int32_t* cache_hits, int32_t* failures);
int nebo_fused_source(const nebo_expr_t* rhs, int32_t reduce_op, char* buf, int32_t size,
                      int32_t* needed);
/* (host only) What nebo_assign would launch for lhs <<= rhs when the tree runs on
 * the compiled column schedule: the kernel source (as nebo_fused_source), its
 * argument struct (args_size bytes at least: 1280), grid and block sizes.
 * NEBO_ERR_UNSUPPORTED when the tree takes another path.  No counterpart in the
 * reference: it lets the tests run the generated kernels' address arithmetic on
 * the host and check every access against the fields' allocations. */
int nebo_fused_column_launch(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost, char* src,
                             int32_t src_size, int32_t* src_needed, void* args, int32_t args_size,
                             int32_t* grid, int32_t* block);

/* Which kernel nebo_assign would pick for this tree: 0 generic evaluator,
 * 1 var-coefficient diffusion, 2 Laplacian+Euler, 3 pointwise copy/scale.
 * Host-only. */
int nebo_assign_plan(const nebo_field_t* lhs, const nebo_expr_t* rhs,
                     int use_ghost, int* kernel_id);

/* Host-only: size of the fused program the generic evaluator runs for rhs (its
 * instruction count, evaluation-stack depth and distinct (field, offset) reads).
 * Fails with NEBO_ERR_TOO_LARGE when the tree exceeds the evaluator's limits
 * (200 instructions, depth 10, 40 reads, 48 constants, 6 mask offsets). */
int nebo_lower_stats(const nebo_expr_t* rhs, int32_t* n_insns, int32_t* max_depth,
                     int32_t* n_loads);

/* masked_assign(mask, lhs, rhs): lhs(p) = rhs(p) for each listed point p
 * (interior-relative ijk triples, HOST array of 3*n_points int32; the library
 * checks every point against the LHS window and the RHS's valid ghost range,
 * then copies the list to the device).  The reference evaluates this on CPU only
 * (NeboLhs.h:164-283, NeboAssignment.h:246-313); this is its GPU counterpart,
 * with one evaluation per point.  The RHS may read the LHS through stencils
 * (the boundary-condition inversion of NeboStencilBuilder.h:590-764 does) as long
 * as no point reads a point that another point writes; such lists, and lists that
 * repeat a point while the RHS reads the LHS, fail with NEBO_ERR_INVALID because
 * their result would depend on the order of the writes.  Reads of edgeless
 * stencils are checked point by point against each operand's allocated block. */
int nebo_masked_assign(const nebo_field_t* lhs, const int32_t* points,
                       int32_t n_points, const nebo_expr_t* rhs, void* stream);

/* ------------------------------------------------------------------------- */
/* reductions (NeboReductions.h:45-240; CPU-only in the reference)           */
/* ------------------------------------------------------------------------- */
enum nebo_reduce_op { NEBO_RED_MAX = 0, NEBO_RED_MIN = 1, NEBO_RED_SUM = 2, NEBO_RED_NORM = 3 };

/* field_max/min/sum/norm (use_ghost=1) and the _interior variants (use_ghost=0)
 * of an expression.  Synchronises `stream` and writes the host scalar *out.
 * max/min are exact; sum/norm accumulate in double-double (error-free
 * transformations) and round once, so they are within one rounding of the
 * exact sum.  An expression that reads no field and no mask has no extent and
 * "infinite" ghosts: with use_ghost=1 its region is unbounded (the reference's
 * loop would walk [-GHOST_MAX, GHOST_MAX)^3) and the call fails with
 * NEBO_ERR_INVALID; with use_ghost=0 the region is empty and *out is the
 * fold's initial value (-inf, +inf, 0, 0). */
int nebo_reduce(const nebo_expr_t* expr, int op, int use_ghost, double* out,
                void* stream);

/* Same, asynchronous: writes one double to device memory out_dev.  Two kernel
 * launches (one partial per block, then a fixed-order fold); the partials live in
 * a buffer the library creates at the first reduction on each stream and keeps,
 * so later calls on that stream allocate nothing. */
int nebo_reduce_async(const nebo_expr_t* expr, int op, int use_ghost,
                      double* out_dev, void* stream);

/* nebo_reduce_async that allocates no device memory, ever.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  (The first
 * call with a new descriptor keeps a small host-side record for its repeats.)  The
 * partials go to a static device array of 32 KB per GPU (the partials of at most
 * 2048 blocks plus the result slot; grids never exceed 2048 blocks, so no input
 * size needs more), shared by every call on that GPU and holding the last call's
 * contents: calls that overlap on one GPU (on different streams, or from
 * different threads) race and must not be issued.  It runs the kernel compiled
 * for the tree's shape when an earlier allocating call compiled it (then its
 * result equals nebo_reduce_async's bit for bit), else the precompiled
 * interpreter kernels (max/min equal, sum/norm within one rounding).  The
 * operands must be on the current device. */
int nebo_reduce_async_noalloc(const nebo_expr_t* expr, int op, int use_ghost, double* out_dev,
                              void* stream);

/* ------------------------------------------------------------------------- */
/* ghost cells                                                               */
/* ------------------------------------------------------------------------- */

/* Periodic copy-fill of every allocated ghost layer of the field from the
 * opposite interior side (the reference leaves ghost filling to its caller;
 * this is the fill the periodic workloads use). */
int nebo_fill_ghosts_periodic(const nebo_field_t* field, void* stream);

/* ------------------------------------------------------------------------- */
/* direct entry points of the hand-written stencil kernels                    */
/* ------------------------------------------------------------------------- */

/* Per-axis coefficients of the staggered operators of build_stencils
 * (spatialops/structured/stencil/StencilBuilder.cpp:60-72). */
typedef struct nebo_axis_coefs {
  double interp[3][2];   /* InterpC2F{X,Y,Z}: {0.5, 0.5}           */
  double grad[3][2];     /* Grad{X,Y,Z}:      {-1/dx, +1/dx}       */
  double div[3][2];      /* Div{X,Y,Z}:       {-1/dx, +1/dx}       */
} nebo_axis_coefs_t;

/* r <<= divX(interpX(k)*gradX(phi)) + divY(..) + divZ(..)     (form 0)
 * r <<= -divX(gradX(phi)*interpX(k)) - divY(..) - divZ(..)    (form 1,
 *       test/StencilTest.cpp:193-195)
 * over r's interior (the ghost inference of the tree gives interior only). */
int nebo_divkgrad(const nebo_field_t* r, const nebo_field_t* k,
                  const nebo_field_t* phi, const nebo_axis_coefs_t* c, int form,
                  void* stream);

/* out <<= phi + s*(divX(gradX(phi)) + divY(gradY(phi)) + divZ(gradZ(phi)))
 * over out's interior: one explicit-Euler heat-equation step with s = dt*alpha. */
int nebo_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                    const nebo_axis_coefs_t* c, double s, void* stream);

/* Tile of the hand-written stencil kernels: a workgroup covers tx (x) by ty (y)
 * columns and marches zchunk planes.  variant 0 (tx = 64) reads every neighbour
 * straight from memory; variants 1 and 2 (tx = 64) stage each plane's rows in
 * LDS, take x neighbours from the adjacent lanes and prefetch 1 or 2 planes
 * ahead; variant 3 (column schedule) puts tx/64 waves side by side in x, each
 * lane computing ty rows of one column with no LDS and no barrier, loads issued
 * two planes ahead; variant 4 is variant 3 computing each x face once (the
 * Laplacian kernel runs variant 3 for it); variants 5 and 6 stack ty/4 resp.
 * ty/2 such waves in y in one workgroup (4 resp. 2 rows per lane), so that a
 * wave's halo rows are its neighbours' own rows, loaded on the same CU; variants
 * 7 and 8 are variant 5 with loads issued 3 resp. 4 planes ahead (the Laplacian
 * kernel runs variant 3 with the same rows per lane for 5-8); variant 9 is 7 with
 * write-back instead of nontemporal stores.  Variants 10-16 (pair columns,
 * div(k grad phi) only; the Laplacian runs its column schedule over the same
 * tile) give each lane two adjacent columns, so that every load and store moves
 * 16 bytes and a wave covers 128 columns (tx a multiple of 128; boxes whose width
 * is not a multiple of tx run the rule-4 tile instead), with rows per lane and
 * planes of prefetch 10: 4, 3; 11: 2, 3; 12: 4, 2; 13: 2, 4; 14: 2, 2; 15: 1, 3;
 * 16: 1, 4.  kernel 0 = the
 * var-coefficient diffusion kernels (nebo_divkgrad, nebo_diffusion_step), 1 =
 * Laplacian+Euler.  Which tile a launch uses:
 *   1. the tile pinned by nebo_stencil_set_tile, for every box;
 *   2. else the tile nebo_stencil_autotune found fastest for a box of the same
 *      dimensions on the same device;
 *   3. else 128 x 8 tiles, variant 7, with the longest z chunk of 512, 256 or
 *      128 planes that gives every CU a workgroup (512^3: 512 planes);
 *   4. else 128 x 8 x 64, variant 11 for kernel 0 on boxes whose width is a
 *      multiple of 128 and variant 7 otherwise, when it gives at least half the
 *      device's CUs a workgroup (256^3);
 *   5. else (small boxes) variant 3 with 64 x 2 tiles (one wave, 2 rows per
 *      lane) and the longest z chunk of 16, 8 or 4 planes that gives the device
 *      two workgroups per CU, else 4 planes.
 * nebo_stencil_set_tile with tx = ty = zchunk = variant = 0 removes the pin.
 * nebo_stencil_get_tile reports the tile last pinned or tuned (the default tile
 * if neither).  The tile changes only the schedule, never the results.  The
 * reference has no counterpart (its CUDA launch shape is fixed at 16x16,
 * NeboLhs.h:413-427). */
int nebo_stencil_tiles(int32_t* n, int32_t* tx, int32_t* ty, int32_t* zchunk,
                       int32_t* variant);                               /* arrays of n */
int nebo_stencil_set_tile(int kernel, int32_t tx, int32_t ty, int32_t zchunk, int32_t variant);
int nebo_stencil_get_tile(int kernel, int32_t* tx, int32_t* ty, int32_t* zchunk, int32_t* variant);
/* The tile a launch of `kernel` over an nx x ny x nz box on `device` would use
 * (rules 1-5 above). */
int nebo_stencil_select_tile(int kernel, int device, int32_t nx, int32_t ny, int32_t nz, int32_t* tx,
                             int32_t* ty, int32_t* zchunk, int32_t* variant);
/* Time every instantiated tile of `kernel` on these operands (r is overwritten on
 * its interior; k is unused for kernel 1), remember the fastest for boxes of
 * these dimensions on this device (removing any pin of that kernel) and report
 * it; times_ms (optional) receives one mean time per tile in nebo_stencil_tiles
 * order.  Synchronises stream. */
int nebo_stencil_autotune(int kernel, const nebo_field_t* r, const nebo_field_t* k,
                          const nebo_field_t* phi, const nebo_axis_coefs_t* c, void* stream,
                          int32_t* tx, int32_t* ty, int32_t* zchunk, int32_t* variant,
                          float* times_ms);

/* Streaming ceiling: r[i] = a[i] + s*b[i] for i < n, the two-read one-write
 * traffic of div(k grad phi) without halos or arithmetic.  No counterpart in the
 * reference: a calibration kernel, timed beside the stencil kernels so their HBM
 * rate is read against what a plain stream reaches on the same blocks.
 * variant 0: 8-byte accesses, nontemporal stores; 1: 16-byte accesses,
 * nontemporal stores; 2: 16-byte accesses, write-back stores; 3: two 16-byte
 * accesses per lane and array, nontemporal stores (1-3 need 16-byte aligned
 * arrays); 4: 16-byte accesses, nontemporal stores, on arrays that start 8 bytes
 * past a 16-byte boundary (as interior rows of a one-ghost field do). */
int nebo_stream_triad(double* r, const double* a, const double* b, double s, int64_t n, int32_t variant,
                      void* stream);

/* ------------------------------------------------------------------------- */
/* batched patches: one persistent launch over a device patch table          */
/* ------------------------------------------------------------------------- */

/* One Uintah-style patch: its own interior cell counts and one ghost layer.
 * a and b are two (n[0]+2)*(n[1]+2)*(n[2]+2) x-fastest blocks of the patch's
 * state; a step reads one and writes the interior of the other (ghost cells are
 * never written: they hold whatever boundary values the caller put there). */
typedef struct nebo_patch {
  double* a;
  double* b;
  int32_t n[3];          /* interior cells per axis (each >= 1)             */
  int32_t pad_;
} nebo_patch_t;

typedef struct nebo_batch nebo_batch_t;   /* opaque */

/* Build a device patch table and the tile schedule for `n` patches, all on the
 * current device.  tile[3] is the (x, y, z) extent of one work item (x*y <= 1024
 * threads), or with y = 0 a flattened item of x consecutive points of a patch's
 * x-y plane (x a multiple of 64 up to 1024) by z planes, which leaves no lane
 * idle on patches narrower than the tile, or with x = 0 the column schedule: one
 * wave per item covering up to 64 columns of a patch (64/nx row groups side by
 * side on patches narrower than 64), y = 1, 2 or 4 rows per lane, z planes per
 * item (items of up to 8 planes load all their planes as they start), x
 * neighbours exchanged between lanes; tile == NULL starts from (64, 4, 16).  Coefficients and s are shared by all patches:
 * out = in + s*((Lx + Ly) + Lz) as in nebo_lapl_euler. */
int nebo_batch_create(const nebo_patch_t* patches, int32_t n,
                      const nebo_axis_coefs_t* c, double s,
                      const int32_t* tile, nebo_batch_t** out);
int nebo_batch_destroy(nebo_batch_t* b);
/* One step on every patch in one persistent launch: a -> b (swap = 0) or
 * b -> a (swap = 1). */
int nebo_batch_step(nebo_batch_t* b, int swap, void* stream);
/* The same step as one launch per patch (the comparison path). */
int nebo_batch_step_per_patch(nebo_batch_t* b, int swap, void* stream);
/* Capture `steps` alternating steps (a->b, b->a, ...) into a hipGraph, built
 * once and replayed by nebo_batch_graph_launch.  stream must not be NULL. */
int nebo_batch_graph_build(nebo_batch_t* b, int32_t steps, void* stream);
int nebo_batch_graph_launch(nebo_batch_t* b, void* stream);
/* Time candidate tiles on alternating steps (a -> b, b -> a, as a time loop
 * runs them; the blocks' contents change), keep the fastest, write it to
 * tile_out (3 ints) and drop any captured graph.  Synchronises stream. */
int nebo_batch_autotune(nebo_batch_t* b, int32_t* tile_out, void* stream);
/* Tile currently used, number of work items and of persistent workgroups. */
int nebo_batch_info(nebo_batch_t* b, int32_t* tile_out, int32_t* n_items,
                    int32_t* n_blocks);

/* ------------------------------------------------------------------------- */
/* adjoint of the var-coefficient diffusion step                              */
/* ------------------------------------------------------------------------- */

/* Forward explicit-Euler step on a periodic field:
 *   fill_ghosts_periodic(phi); out = phi + dt*div(interp(k)*grad(phi))
 * (interior), then fill_ghosts_periodic(out).  k must have periodic ghosts. */
int nebo_diffusion_step(const nebo_field_t* out, const nebo_field_t* phi,
                        const nebo_field_t* k, const nebo_axis_coefs_t* c,
                        double dt, void* stream);

/* Adjoint of nebo_diffusion_step (transposed Div∘Gradient with variable
 * coefficient).  Given lam = dL/d(out) (periodic ghosts filled) and the step's
 * input phi (periodic ghosts filled):
 *   lam_out = lam + dt*A(k)^T lam        (dL/d phi)
 *   gk     += dt * I^T( (D^T lam) .* (G phi) )   (dL/dk, interior, accumulated)
 * lam_out gets periodic ghosts filled. */
int nebo_diffusion_adjoint_step(const nebo_field_t* lam_out, double* gk,
                                const nebo_field_t* lam, const nebo_field_t* phi,
                                const nebo_field_t* k, const nebo_axis_coefs_t* c,
                                double dt, void* stream);

/* Whole trajectory: n_steps forward steps from phi0 (n_steps a multiple of
 * `every`), keeping the state after every `every` steps: ckpt holds
 * n_steps/every + 1 glob blocks of phi0's shape back to back (block j = state
 * after j*every steps, block 0 = phi0 with periodic ghosts); the final state
 * also goes to phiT. */
int nebo_diffusion_forward(const nebo_field_t* phiT, const nebo_field_t* phi0,
                           const nebo_field_t* k, const nebo_axis_coefs_t* c,
                           double dt, int32_t n_steps, int32_t every,
                           double* ckpt, void* stream);

/* Reverse sweep with checkpoint/recompute: lamT = dL/d phiT on entry (its
 * ghosts are filled here).  Each segment of `every` steps is recomputed from its
 * checkpoint, then swept backwards with the adjoint step.  On return
 * lam0 = dL/d phi0 and gk (k's shape) = dL/dk on interior cells, ghosts zero.
 * work must hold every+1 glob blocks of phi0's shape. */
int nebo_diffusion_backward(const nebo_field_t* lam0, double* gk,
                            const nebo_field_t* lamT, const nebo_field_t* k,
                            const nebo_axis_coefs_t* c, double dt,
                            int32_t n_steps, int32_t every, const double* ckpt,
                            double* work, void* stream);

/* ------------------------------------------------------------------------- */
/* scalar transport chain: upwinded convection + var-coefficient diffusion,   */
/* SSP-RK3, periodic, and its adjoint                                         */
/* ------------------------------------------------------------------------- */

/* Schedule of the transport stage kernel: 0 chosen by box size (the default),
 * 1 direct (64 x 4 columns of cells per workgroup, every neighbour loaded),
 * 2 column tiles of 128 x 8 cells (each lane a column of 4 rows; x neighbours by
 * lane exchange, y neighbours from the lane's rows) with loads 2 planes ahead,
 * 3 the same with loads 1 plane ahead.  By box size: column tiles with loads 1
 * plane ahead and the longest z chunk of 512, 256, 128 or 64 planes that gives
 * every CU a workgroup, else the direct schedule.  zchunk: planes per workgroup, 0 for the schedule's
 * default (16 for the direct one).  The adjoint stage always runs the direct
 * schedule (a zchunk set with schedule 1 applies to it).  Changes only the
 * schedule, never the results.  nebo_transport_schedule reports the schedule
 * and chunk a stage over `field`'s interior would use.  No reference
 * counterpart. */
int nebo_transport_tune(int32_t schedule, int32_t zchunk);
int nebo_transport_schedule(const nebo_field_t* field, int32_t* schedule, int32_t* zchunk);

/* One SSP-RK3 stage of R(s) = div(interp(k)*grad(s)) - div(u*upwind(s)), where
 * upwind(s) on a face is cond(u > 0, s of the cell below)(s of the cell above):
 *   stage 1: out = phin + dt*R(phin)                  (s unused, may be NULL)
 *   stage 2: out = 0.75*phin + 0.25*(s + dt*R(s))
 *   stage 3: out = (1/3)*phin + (2/3)*(s + dt*R(s))
 * over the interior, then out's periodic ghosts.  u points to the three face
 * velocity fields (x, y, z faces).  Every operand has one periodic ghost layer
 * and the same window.  Bit-identical to nebo_assign of the fused stage
 * expression (built from the reference's Divergence/Interpolant/Gradient
 * stencils and cond, NeboStencilBuilder.h:134-148, NeboCond.h:3801-4465)
 * followed by nebo_fill_ghosts_periodic. */
int nebo_transport_stage(const nebo_field_t* out, const nebo_field_t* phin,
                         const nebo_field_t* s, const nebo_field_t* k,
                         const nebo_field_t* u, const nebo_axis_coefs_t* c, double dt,
                         int stage, void* stream);
/* Transpose of one stage for adjoint a (periodic ghosts), with X = a + dt*R_s^T a:
 *   stage 3: out = (2/3)*X,        acc_out = (1/3)*a,          gk += (2/3)*dt*R_k(s)^T a
 *   stage 2: out = 0.25*X,         acc_out = acc_in + 0.75*a,  gk += 0.25*dt*R_k(s)^T a
 *   stage 1: out = acc_in + X,                                  gk += dt*R_k(s)^T a
 * (R_s: R as a linear map of s; R_k(s): its diffusion part as a linear map of k.)
 * acc_in / acc_out / gk are blocks of the field's shape (interior cells used);
 * out's periodic ghosts are filled.  No reference counterpart (the reference has
 * no adjoint). */
int nebo_transport_adjoint_stage(const nebo_field_t* out, double* acc_out, double* gk,
                                 const nebo_field_t* a, const double* acc_in,
                                 const nebo_field_t* s, const nebo_field_t* k,
                                 const nebo_field_t* u, const nebo_axis_coefs_t* c,
                                 double dt, int stage, void* stream);
/* n_steps RK3 steps from phi0 (its periodic ghosts are filled in the first
 * checkpoint), keeping phi0 and every `every`-th state: ckpt holds
 * n_steps/every + 1 blocks.  phiT receives the final state. */
int nebo_transport_forward(const nebo_field_t* phiT, const nebo_field_t* phi0,
                           const nebo_field_t* k, const nebo_field_t* u,
                           const nebo_axis_coefs_t* c, double dt, int32_t n_steps,
                           int32_t every, double* ckpt, void* stream);
/* Reverse sweep for lamT = dL/d phiT: each segment is recomputed from its
 * checkpoint (stage states kept), then swept back stage by stage.  On return
 * lam0 = dL/d phi0 and gk = dL/dk on interior cells (ghosts zero).  work holds
 * 3*every + 5 blocks. */
int nebo_transport_backward(const nebo_field_t* lam0, double* gk, const nebo_field_t* lamT,
                            const nebo_field_t* k, const nebo_field_t* u,
                            const nebo_axis_coefs_t* c, double dt, int32_t n_steps,
                            int32_t every, const double* ckpt, double* work, void* stream);

/* ------------------------------------------------------------------------- */
/* particles and point fields                                                */
/* ------------------------------------------------------------------------- */

/* Constructor arguments of ParticleToCell / CellToParticle
 * (spatialops/particles/ParticleOperators.h:151-166): mesh spacing and the
 * coordinate of the first interior cell centre per axis.  A spacing <= 0 means
 * the default (spacing 1, origin 0), as there. */
typedef struct nebo_particle_geom {
  double dx, xlo, dy, ylo, dz, zlo;
} nebo_particle_geom_t;

/* ParticleToCell<CellField>::apply_to_field (ParticleOperators.h:184-270): dest
 * <<= 0, then each particle p (a cube of edge psize[p] centred at (px, py, pz))
 * adds src[p] * (its volume fraction in the cell) to every cell of dest it
 * overlaps.  px / py / pz may be NULL for an axis the particles do not resolve.
 * Every cell receives its contributions in particle order, so the sums are those
 * of the reference's serial loop bit for bit: the contributions, written in the
 * serial loop's order, are stably sorted by cell and summed per cell.  A particle whose box begins a cell or
 * more below the window on some axis contributes nothing (the reference's
 * unsigned index conversion wraps there); `skipped` (device int, may be NULL)
 * receives how many did.  All particle arrays hold np doubles on dest's GPU. */
int nebo_particle_to_cell(const nebo_field_t* dest, const double* src, const double* px,
                          const double* py, const double* pz, const double* psize, int64_t np,
                          const nebo_particle_geom_t* geom, int32_t* skipped, void* stream);

/* CellToParticle<CellField>::apply_to_field (ParticleOperators.h:305-395): dest[p]
 * = sum over the cells the particle overlaps, in k, j, i order, of src(i,j,k) *
 * (its volume fraction there); one thread per particle, same operations in the
 * same order as the reference.  Skipped particles (as above) get 0. */
int nebo_cell_to_particle(double* dest, const nebo_field_t* src, const double* px,
                          const double* py, const double* pz, const double* psize, int64_t np,
                          const nebo_particle_geom_t* geom, int32_t* skipped, void* stream);

/* The index list of PointToField / FieldToPoint (spatialops/pointfield/
 * PointOperators.h:40-160): flat indices into a field's block, uploaded once to
 * `device`.  Where an index repeats, PointToField keeps the last point (the
 * reference's loop writes in order).  Indices are taken from HOST memory. */
typedef struct nebo_point_map nebo_point_map_t;
int nebo_point_map_create(const int64_t* indices, int64_t n, int device, nebo_point_map_t** out);
int nebo_point_map_destroy(nebo_point_map_t* map);
/* PointToField::apply_to_field: dest.values[index[i]] = src[i] for i < n. */
int nebo_point_to_field(const nebo_point_map_t* map, const nebo_field_t* dest, const double* src,
                        void* stream);
/* FieldToPoint::apply_to_field: dest[i] = src.values[index[i]] for i < n. */
int nebo_field_to_point(const nebo_point_map_t* map, double* dest, const nebo_field_t* src,
                        void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NEBO_HIP_H */
