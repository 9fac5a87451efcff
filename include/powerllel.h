/*
 * PowerLLEL-MI355X: C ABI of the GPU time-step hot path of the channel-flow
 * solver (FFT pressure Poisson solve + RK2 projection step) for gfx950.
 *
 * Plain C types only (no torch, no HIP types in signatures).  Streams are
 * passed as `void *` holding a hipStream_t (NULL = the null stream).
 *
 * Array conventions are the reference's (xiejb6/PowerLLEL_Channel):
 *  - a haloed field f(0:nx+1, 0:ny+1, 0:nz+1) is one contiguous double array,
 *    x fastest: Fortran f(i,j,k) is f[i + (nx+2)*(j + (ny+2)*k)]
 *    (src/mod_variables.f90:14-44, src/src_c/mod_poisson_solver.c:187-188);
 *  - dzf_global has nz+2 entries, dzf_global[0] and dzf_global[nz+1] are the
 *    ghost cell heights (src/mod_mesh.f90:35,156-157).
 *
 * Error behaviour: the three drop-in entry points are `void`, like the
 * reference's; on an error they print a message and abort the process (the
 * reference aborts through MPI_Abort).  Every other entry point returns
 * PLL_OK (0) or a negative PLL_E* code; pll_last_error() describes the last
 * failure of the calling thread.
 */
#ifndef POWERLLEL_H
#define POWERLLEL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PLL_OK = 0,
    PLL_EINVAL = -1,   /* bad argument (sizes, BC strings, NULL pointers)  */
    PLL_ENOMEM = -2,   /* device or host allocation failed                 */
    PLL_EHIP = -3,     /* a HIP runtime call or kernel launch failed       */
    PLL_ESTATE = -4,   /* called out of order (execute before init, ...)   */
    PLL_EUNSUP = -5    /* a configuration the library does not implement   */
};

/* Description of the last error on this thread ("" if none). */
const char *pll_last_error(void);

/* ------------------------------------------------------------------------
 * Drop-in PPE solver: the reference's three C entry points, same
 * signatures and semantics, one global instance.
 * ---------------------------------------------------------------------- */

/* Replaces init_poisson_solver, src/src_c/mod_poisson_solver.c:178-309
 * (bound from Fortran at src/mod_poissonSolver.f90:61-72, called at
 * src/main.f90:80-81).  BC strings: bctype_x = bctype_y = "PP";
 * bctype_z in {"NN","DD","ND","DN"}.  neighbor_xyz[2][4], [2][5] (the
 * z-pencil's bottom/top ranks) must be MPI_PROC_NULL (any negative value):
 * the whole domain lives on one GPU. */
void init_poisson_solver(int nx_global, int ny_global, int nz_global,
                         double dx, double dy, double *dzf_global,
                         char bctype_x[], char bctype_y[], char bctype_z[],
                         int neighbor_xyz[3][6]);

/* Replaces execute_poisson_solver, src/src_c/mod_poisson_solver.c:468-605
 * (called every step at src/main.f90:177).  `var` is the haloed p array
 * holding the right-hand side in its interior; on return the interior holds
 * the solution and the halos are untouched.  `var` may be host memory
 * (copied to and from the GPU) or device memory (solved in place). */
void execute_poisson_solver(double *var);

/* Replaces free_poisson_solver, src/src_c/mod_poisson_solver.c:607-632. */
void free_poisson_solver(void);

/* Status-returning forms of the same three calls (no abort). */
int init_poisson_solver_status(int nx_global, int ny_global, int nz_global,
                               double dx, double dy, const double *dzf_global,
                               const char *bctype_x, const char *bctype_y,
                               const char *bctype_z, const int neighbor_xyz[3][6]);
int execute_poisson_solver_status(double *var);
/* Device-pointer form of execute_poisson_solver on `stream`
 * (asynchronous: the caller synchronises). */
int execute_poisson_solver_dev(double *d_var, void *stream);
int free_poisson_solver_status(void);

/* ------------------------------------------------------------------------
 * Handle API of the same solver (several independent instances).
 * ---------------------------------------------------------------------- */
typedef struct pll_ppe pll_ppe;

/* Same meaning as init_poisson_solver's arguments (one rank, walls in z). */
int pll_ppe_create(pll_ppe **out, int nx, int ny, int nz, double dx, double dy,
                   const double *dzf_global, const char *bctype_x,
                   const char *bctype_y, const char *bctype_z);
/* Solve in place on the device haloed array d_var (interior in/out, halos
 * untouched), asynchronously on `stream`. */
int pll_ppe_solve(pll_ppe *h, double *d_var, void *stream);
/* Stage timing with HIP events recorded on the solve's stream (no
 * counterpart in the reference, which times stages with GPTL,
 * mod_poisson_solver.c:480-602).  Turning timing on (re)starts the record;
 * stages are x-fwd, y-fwd, z-solve, y-bwd, x-bwd.  pll_ppe_stage_times
 * gives the last timed solve, pll_ppe_stage_times_mean the mean over the
 * last (up to 128) timed solves and their count.  Both synchronise. */
int pll_ppe_set_timing(pll_ppe *h, int on);
int pll_ppe_stage_times(pll_ppe *h, float ms[5]);
int pll_ppe_stage_times_mean(pll_ppe *h, double ms[5], int *nsolves);
/* Refined form (no counterpart in the reference, whose solver is FP64
 * only): the transforms, the spectral array and the z solves run in FP32
 * storage (z recurrence in FP64 registers); then `sweeps` rounds of
 * r = f - L_h p computed in FP64, an FP32 solve L c = r and p += c.  The
 * operator, gauge and in-place semantics are those of pll_ppe_solve.  The
 * first call allocates 2 N floats + N doubles of work space. */
int pll_ppe_solve_refined(pll_ppe *h, double *d_var, int sweeps, void *stream);
/* Make pll_ppe_solve (and so the drop-in execute_poisson_solver) use the
 * refined form with `sweeps` sweeps; -1 restores the FP64 solve. */
int pll_ppe_set_refinement(pll_ppe *h, int sweeps);
/* Same switch for the drop-in global instance (after init_poisson_solver). */
int set_poisson_solver_refinement(int sweeps);
/* Refined form to a tolerance (no counterpart in the reference): sweeps
 * until the estimated relative max-norm error of p is <= tol, at most
 * max_sweeps (>= 1).  After sweep k the estimate is |c_1|/|p| for k = 1 and
 * min(1, 10 |c_k|/|c_{k-1}|) |c_k|/|p| for k > 1, c_k the k-th correction.
 * Reads two norms back per sweep (one stream synchronisation each).  The
 * sweeps made and the final estimate go to *sweeps_done / *err_est (either
 * may be NULL).  Work space: that of pll_ppe_solve_refined plus
 * 2 + 2 ny nz doubles. */
int pll_ppe_solve_refined_tol(pll_ppe *h, double *d_var, double tol,
                              int max_sweeps, int *sweeps_done,
                              double *err_est, void *stream);
/* Make pll_ppe_solve use the tolerance form (pll_ppe_set_refinement
 * switches it off), and query the sweeps and estimate of its last solve. */
int pll_ppe_set_refinement_tol(pll_ppe *h, double tol, int max_sweeps);
int pll_ppe_last_refinement(const pll_ppe *h, int *sweeps, double *err_est);
/* Device bytes held by the handle. */
size_t pll_ppe_device_bytes(const pll_ppe *h);
int pll_ppe_destroy(pll_ppe *h);

/* ------------------------------------------------------------------------
 * Step-level entries: the reference's per-step Fortran routines with the
 * same array conventions, on device arrays, asynchronous on `stream`.
 * ---------------------------------------------------------------------- */
typedef struct pll_channel pll_channel;

/* Grid and physics: dzf_global as above; re_inv = nu
 * (src/mod_parameters.f90:187); u_crf the convecting-frame velocity;
 * is_forced / vel_force as the &FORCE namelist (src/mod_calcVel.f90:455-509).
 * vel_force is used as given (the caller applies transform2CRF's shift,
 * src/mod_calcVel.f90:47, if u_crf != 0). */
int pll_channel_create(pll_channel **out, int nx, int ny, int nz, double lx,
                       double ly, double lz, const double *dzf_global,
                       double re_inv, double dt, double u_crf,
                       const int is_forced[3], const double vel_force[3]);
int pll_channel_destroy(pll_channel *c);

/* timeIntVelRK1_kernel, src/mod_calcVel.f90:52-146 (interior). */
int pll_channel_rk1(pll_channel *c, const double *u, const double *v,
                    const double *w, double *u1, double *v1, double *w1,
                    void *stream);
/* timeIntVelRK2_kernel, src/mod_calcVel.f90:240-334 (interior, u,v,w in place). */
int pll_channel_rk2(pll_channel *c, double *u, double *v, double *w,
                    const double *u1, const double *v1, const double *w1,
                    void *stream);
/* calcRHS, src/mod_calcRHS.f90:11-37 (writes the interior of p). */
int pll_channel_calc_rhs(pll_channel *c, const double *u, const double *v,
                         const double *w, double *p, void *stream);
/* correctVel, src/mod_calcVel.f90:428-453. */
int pll_channel_correct_vel(pll_channel *c, const double *p, double *u,
                            double *v, double *w, void *stream);
/* forceVel + Mean_h, src/mod_calcVel.f90:455-509, src/mod_utils.f90:105-131. */
int pll_channel_force_vel(pll_channel *c, double *u, double *v, double *w,
                          void *stream);
/* updateBoundVel at one rank, src/mod_updateBound.f90:30-108. */
int pll_channel_bc_vel(pll_channel *c, double *u, double *v, double *w,
                       void *stream);
/* updateBoundP at one rank, src/mod_updateBound.f90:186-239. */
int pll_channel_bc_p(pll_channel *c, double *p, void *stream);
/* Mean_h, src/mod_utils.f90:105-131 (synchronises, result on the host). */
int pll_channel_mean_h(pll_channel *c, const double *q, double *mean,
                       void *stream);
/* calcMaxDiv, src/mod_utils.f90:257-318: max |div u| and its 1-based
 * (i,j,k) (first in k,j,i order among equal maxima).  Synchronises. */
int pll_channel_max_div(pll_channel *c, const double *u, const double *v,
                        const double *w, double *value, int ijk[3],
                        void *stream);
/* calcMaxCFL, src/mod_utils.f90:193-255: max over the interior of
 * dt (|u_i + u_{i-1}|/2 dx_inv + |v_j + v_{j-1}|/2 dy_inv + |w_k + w_{k-1}|/2 dzf_inv(k))
 * and its 1-based (i,j,k) (first in k,j,i order among equal maxima; NaN
 * cells are skipped, as by the reference's comparisons).  Synchronises. */
int pll_channel_max_cfl(pll_channel *c, const double *u, const double *v,
                        const double *w, double *value, int ijk[3],
                        void *stream);
/* checkNaN, src/mod_utils.f90:352-394: *has_nan = 1 if any element of the
 * haloed field is NaN.  Synchronises. */
int pll_channel_check_nan(pll_channel *c, const double *q, int *has_nan,
                          void *stream);
/* One whole time step, src/main.f90:137-208: RK1, BC, RK2, BC, calcRHS,
 * PPE (ppe must match the grid), updateBoundP, correctVel, forceVel, BC.
 * With the FP64 solve, calcRHS is formed inside the solve's x transform
 * (same arithmetic).  Forcing of u and/or v takes Mean_h from the RK2 output
 * and shifts inside correctVel: the x (y) pressure-gradient correction sums
 * to zero over a periodic row, so this is the same mean up to summation
 * roundoff.  A forced w, or the refined solve, uses the separate passes.
 * (u1, v1, w1) is work space for the RK1 stage: RK2 reads its boundary
 * values from its interior, so the step leaves its boundary fill (ghost
 * cells, w1's wall row) unmade. */
int pll_channel_step(pll_channel *c, pll_ppe *ppe, double *u, double *v,
                     double *w, double *u1, double *v1, double *w1, double *p,
                     void *stream);
/* nsteps calls of pll_channel_step (src/main.f90:131-212 without the
 * per-step checks and output), with the same results bit for bit.  With the
 * FP64 solve, w not forced and a large grid (at least 2048 RK tiles of
 * 64 x 4 x 64 cells), each step's projection (correctVel, forcing, boundary
 * fill) is formed inside the next step's RK1 and RK2 as they read the
 * velocity, which saves one pass over the fields per step; the run works in
 * place and allocates no device memory. */
int pll_channel_run(pll_channel *c, pll_ppe *ppe, int nsteps, double *u,
                    double *v, double *w, double *u1, double *v1, double *w1,
                    double *p, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, hosts that forbid
 * hipMalloc): calcMaxDiv, src/mod_utils.f90:257-318, with the results of
 * pll_channel_max_div, but without a channel handle and without any device
 * allocation.  u, v, w are haloed device fields, dzf_global a device array
 * of nz+2 cell heights, dx = lx/nx and dy = ly/ny the cell widths.  Its work
 * space (1/dzf and one partial maximum per (j,k) row) is a static device
 * array of PLL_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends:
 * PLL_EINVAL unless ny*nz <= PLL_NOALLOC_MAX_ROWS and nz+2 <=
 * PLL_NOALLOC_MAX_NZ2.  Calls on one GPU share that array, so calls that
 * overlap on one GPU race.  Synchronises `stream`. */
#define PLL_NOALLOC_MAX_ROWS 524288
#define PLL_NOALLOC_MAX_NZ2 8192
#define PLL_NOALLOC_SCRATCH_BYTES (8 * PLL_NOALLOC_MAX_NZ2 + 16 * PLL_NOALLOC_MAX_ROWS + 256)
int pll_max_div_noalloc(const double *u, const double *v, const double *w,
                        int nx, int ny, int nz, double dx, double dy,
                        const double *dzf_global, double *value, int ijk[3],
                        void *stream);

/* ------------------------------------------------------------------------
 * Adjoint of the projection step (no counterpart in the reference).
 * Every array is a haloed device field like u.
 * ---------------------------------------------------------------------- */
/* pll_channel_step with a body force (bu, bv, bw) added to the momentum
 * right-hand side of both RK stages (interior values are read). */
int pll_channel_step_body(pll_channel *c, pll_ppe *ppe, double *u, double *v,
                          double *w, double *u1, double *v1, double *w1,
                          double *p, const double *bu, const double *bv,
                          const double *bw, void *stream);
/* Wall shear of src/mod_monitor.f90:225-258 (calcAndWriteSkf), averaged over
 * the two walls: re_inv (mean_ij u(i,j,kw) + u_crf) / (dzf(kw)/2).
 * Synchronises. */
int pll_channel_wall_shear(pll_channel *c, const double *u, double *tau,
                           void *stream);
/* calcAndWriteSkf, src/mod_monitor.f90:225-258, at the bottom (k = 1) and
 * top (k = nz) wall: out = {u_tau, cf} bottom, then {u_tau, cf} top, with
 * tau_w = re_inv (mean_ij u(i,j,kw) + u_crf) / (dzf(kw)/2), u_tau =
 * sqrt(tau_w), cf = tau_w / (u_ref^2/2) (the records of monitor_skf_z1.out
 * and monitor_skf_z2.out).  Synchronises. */
int pll_channel_skin_friction(pll_channel *c, const double *u, double u_ref,
                              double out[4], void *stream);
/* extractAndWriteProbePoint, src/mod_monitor.f90:146-167: out = {u + u_crf,
 * v, w, p} at the 1-based interior index (i,j,k) (probe_ijk); PLL_EINVAL for a
 * point outside the domain.  Synchronises. */
int pll_channel_probe(pll_channel *c, const double *u, const double *v,
                      const double *w, const double *p, int i, int j, int k,
                      double out[4], void *stream);
/* Transposes of the step's linear(ised) pieces (each exact for the map the
 * forward kernel applies):
 *  bc_vel_adj / bc_p_adj: ghost adjoints added to the interior values the
 *    fills copy or mirror, ghosts cleared (in place);
 *  force_vel_adj: l -= dzf/lz * sum(l) / (nx ny) for forced components;
 *  correct_vel_adj: lp := (pressure-gradient term)^T (lu, lv, lw), every cell;
 *  calc_rhs_adj: (lu, lv, lw) += calcRHS^T lr (lr read on the interior);
 *  rk_jvp: o = J(q) dq on the interior, J the Jacobian of the momentum
 *    right-hand side of timeIntVelRK1/2;
 *  rk_vjp: o = beta o + a0 l (interior) + alpha J(q)^T l on every cell. */
int pll_channel_bc_vel_adj(pll_channel *c, double *lu, double *lv, double *lw,
                           void *stream);
int pll_channel_bc_p_adj(pll_channel *c, double *lp, void *stream);
int pll_channel_force_vel_adj(pll_channel *c, double *lu, double *lv,
                              double *lw, void *stream);
int pll_channel_correct_vel_adj(pll_channel *c, const double *lu,
                                const double *lv, const double *lw, double *lp,
                                void *stream);
int pll_channel_calc_rhs_adj(pll_channel *c, const double *lr, double *lu,
                             double *lv, double *lw, void *stream);
int pll_channel_rk_jvp(pll_channel *c, const double *qu, const double *qv,
                       const double *qw, const double *du, const double *dv,
                       const double *dw, double *ou, double *ov, double *ow,
                       void *stream);
int pll_channel_rk_vjp(pll_channel *c, const double *qu, const double *qv,
                       const double *qw, const double *lu, const double *lv,
                       const double *lw, double *ou, double *ov, double *ow,
                       double alpha, double a0, double beta, void *stream);
/* Transposed PPE solve (in place, interior): the same transforms with every
 * z system transposed; the gauge-pinned (0,0) system transposes to the same
 * pinned form. */
int pll_ppe_solve_adjoint(pll_ppe *h, double *d_var, void *stream);

/* Gradient of J = (1/nsteps) sum_{n=1..nsteps} wall_shear(u^n) with respect
 * to the body force, over a window of nsteps steps from (u0, v0, w0) (boundary
 * fills applied), with checkpoints every ckpt_every steps and the RK1 stage
 * recomputed in the reverse sweep.  ppe must use the FP64 solve.
 * forward: runs the window, copies the body force, returns J (synchronises).
 * backward: writes dJ/d(bu, bv, bw) (interior; halos 0). */
typedef struct pll_adjoint pll_adjoint;
int pll_adjoint_create(pll_adjoint **out, pll_channel *c, pll_ppe *ppe,
                       int nsteps, int ckpt_every);
int pll_adjoint_forward(pll_adjoint *a, const double *u0, const double *v0,
                        const double *w0, const double *bu, const double *bv,
                        const double *bw, double *J, void *stream);
int pll_adjoint_backward(pll_adjoint *a, double *gu, double *gv, double *gw,
                         void *stream);
size_t pll_adjoint_device_bytes(const pll_adjoint *a);
int pll_adjoint_destroy(pll_adjoint *a);

#ifdef __cplusplus
}
#endif
#endif
