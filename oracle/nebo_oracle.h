/*
 * nebo_oracle.h -- CPU restatement of the reference's Nebo evaluation semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The engine never calls it.
 *
 * Descriptors are the POD types of include/nebo_hip.h, with every `values` /
 * `bits` pointer a HOST pointer.  Each function names the reference lines it
 * restates; see the parity notes at the top of nebo_oracle.c.
 */
#ifndef NEBO_ORACLE_H
#define NEBO_ORACLE_H

#include "../include/nebo_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ghost { int32_t minus[3]; int32_t plus[3]; } oracle_ghost_t;

/* ghosts_with_bc / ghosts_without_bc of the node tree (NeboRhs.h:52-54,183-188;
 * NeboStencils.h:311-317; NeboCond.h:86-91,284-289; NeboMask.h ghosts). */
int oracle_node_ghosts(const nebo_expr_t* e, int node, int with_bc, oracle_ghost_t* g);

/* calculate_actual_ghost + assign limits (NeboBasic.h:61-100, NeboLhs.h:67-88). */
int oracle_assign_limits(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                         nebo_limits_t* lim, int* empty);

/* Value of node at (x,y,z) in the frame of `frame_field`'s valid-ghost-excluded
 * origin (NeboConstField SeqWalk eval, NeboRhs.h:271-285). */
/* boost::math::erf_inv / erfc_inv restated from libm erf/erfc (Newton from
 * Winitzki's approximation); +-inf at the end points, NaN outside the domain. */
double oracle_inv_erf(double y);
double oracle_inv_erfc(double q);
double oracle_eval(const nebo_expr_t* e, int node, int x, int y, int z);

/* lhs <<= rhs, serial z,y,x walk (NeboLhs.h:315-332, 611-623). */
int oracle_assign(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost);

/* lhs <<= rhs with the reference's thread-parallel split (NeboLhs.h:335-393,
 * NeboBasic.h:112-191): one task per partition, partitions along z first. */
int oracle_assign_threads(const nebo_field_t* lhs, const nebo_expr_t* rhs, int use_ghost,
                          int nthreads);

/* masked_assign: lhs(p) = rhs(p) for each point.  passes = 2 restates the
 * CPU-only build, which runs the point loop twice (NeboLhs.h:172-190, 254-275). */
int oracle_masked_assign(const nebo_field_t* lhs, const int32_t* pts, int n,
                         const nebo_expr_t* rhs, int passes);

/* nebo_fold (NeboReductions.h:45-97) with max/min/sum and norm = sqrt(sum(pow(x,2))). */
int oracle_reduce(const nebo_expr_t* e, int op, int use_ghost, double* out);

/* NeboBoundaryConditionBuilder::operator()(mask, phi, gamma, minus) on the CPU:
 * phi is set at x = g - shift for each mask point g so that the operator `op`
 * (phi's location -> gamma's location; points in definition order) applied to
 * phi gives gamma(g) there.  gamma is an expression evaluated in the gamma frame;
 * fd_axis >= 0 marks a same-location FD operator along that axis. */
int oracle_bc_apply(const nebo_field_t* phi, const nebo_expr_t* gamma,
                    const nebo_stencil_t* op, const int32_t phi_off[3],
                    const int32_t gamma_off[3], int fd_axis, const int32_t* pts, int n,
                    int minus);

/* ParticleToCell / CellToParticle::apply_to_field and the point-field
 * operators, serial loops in the reference's order (host arrays). */
int oracle_particle_to_cell(const nebo_field_t* dest, const double* src, const double* px,
                            const double* py, const double* pz, const double* psize, int64_t np,
                            const nebo_particle_geom_t* geom);
int oracle_cell_to_particle(double* dest, const nebo_field_t* src, const double* px,
                            const double* py, const double* pz, const double* psize, int64_t np,
                            const nebo_particle_geom_t* geom);
void oracle_point_to_field(double* field_values, const double* src, const int64_t* idx, int64_t n);
void oracle_field_to_point(double* dest, const double* field_values, const int64_t* idx, int64_t n);

/* Periodic copy-fill of all allocated ghost layers (x, then y, then z sweep). */
void oracle_fill_ghosts_periodic(const nebo_field_t* f);

/* Bit field of a point list over the mask's glob block (BitField.h:291-362). */
void oracle_mask_bits(const nebo_mask_t* m, const int32_t* pts, int n, uint32_t* words);

/* Closed forms (per-cell formulas, written independently of the tree walker). */
void oracle_divkgrad(const nebo_field_t* r, const nebo_field_t* k, const nebo_field_t* phi,
                     const nebo_axis_coefs_t* c, int form);
void oracle_divkgrad_threads(const nebo_field_t* r, const nebo_field_t* k,
                             const nebo_field_t* phi, const nebo_axis_coefs_t* c, int form,
                             int nthreads);
void oracle_lapl_euler(const nebo_field_t* out, const nebo_field_t* phi,
                       const nebo_axis_coefs_t* c, double s);

/* Adjoint of the periodic var-coefficient diffusion step, built from separate
 * operator applications (grad, interp, div and their transposes). */
void oracle_diffusion_step(const nebo_field_t* out, const nebo_field_t* phi,
                           const nebo_field_t* k, const nebo_axis_coefs_t* c, double dt);
void oracle_diffusion_adjoint_step(const nebo_field_t* lam_out, double* gk,
                                   const nebo_field_t* lam, const nebo_field_t* phi,
                                   const nebo_field_t* k, const nebo_axis_coefs_t* c,
                                   double dt);

/* Scalar transport (upwinded convection by the sign of a face velocity plus
 * var-coefficient diffusion), periodic, one ghost layer: one SSP-RK3 stage
 * (stage 1: phin + dt*R(phin); 2: 0.75*phin + 0.25*(s + dt*R(s));
 * 3: (1/3)*phin + (2/3)*(s + dt*R(s))) and its transpose.  u points to the three
 * face-velocity fields (x, y, z).  Both fill the output's periodic ghosts. */
void oracle_transport_stage(const nebo_field_t* out, const nebo_field_t* phin,
                            const nebo_field_t* s, const nebo_field_t* k, const nebo_field_t* u,
                            const nebo_axis_coefs_t* c, double dt, int stage);
/* oracle_transport_stage with the interior split into z slabs over nthreads threads
 * (bit-identical to the serial stage). */
void oracle_transport_stage_threads(const nebo_field_t* out, const nebo_field_t* phin,
                                    const nebo_field_t* s, const nebo_field_t* k, const nebo_field_t* u,
                                    const nebo_axis_coefs_t* c, double dt, int stage, int nthreads);
void oracle_transport_adjoint_stage(const nebo_field_t* out, double* acc_out, double* gk,
                                    const nebo_field_t* a, const double* acc_in,
                                    const nebo_field_t* s, const nebo_field_t* k,
                                    const nebo_field_t* u, const nebo_axis_coefs_t* c, double dt,
                                    int stage);

/* Deterministic inputs shared by CPU and GPU runs. */
uint64_t oracle_splitmix64(uint64_t* state);
void oracle_fill_uniform(double* dst, size_t n, uint64_t seed, double lo, double hi);
void oracle_fill_normal(double* dst, size_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
