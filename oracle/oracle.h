/*
 * Oracle: a CPU restatement of PowerLLEL_Channel's time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  Nothing in powerllel_channel_amd/ links it.
 *
 * What it restates (reference = xiejb6/PowerLLEL_Channel):
 *   - PPE solver: src/src_c/mod_poisson_solver.c:40-103 (set_trid_coeff),
 *     :178-309 (init_poisson_solver), :311-359 (ssolve_trid),
 *     :468-605 (execute_poisson_solver);
 *     src/src_c/mod_fft_fftw.inc:1-33,100-198 (R2HC/HC2R kinds, norm factor,
 *     get_eigen_values).  FFTW (third party, not vendored, >= 3.3.4) is
 *     replaced by an in-file mixed-radix FFT that computes the same
 *     halfcomplex transforms (FFTW_R2HC / FFTW_HC2R definitions).
 *   - Mesh: src/mod_mesh.f90:26-84 (initMesh), :120-162 (initMeshByFunc).
 *   - Step: src/mod_calcVel.f90:26-146 (transform2CRF, RK1 kernel),
 *     :240-334 (RK2 kernel), :428-509 (correctVel, forceVel);
 *     src/mod_calcRHS.f90:11-37; src/mod_utils.f90:105-131 (Mean_h),
 *     :257-318 (calcMaxDiv); src/mod_updateBound.f90:30-108,186-362,364-396
 *     (boundary fills at one rank); src/main.f90:131-212 (step order);
 *     src/mod_monitor.f90:225-258 (wall shear stress).
 *
 * Array convention (the reference's Fortran arrays seen from C): a haloed
 * field f(0:nx+1, 0:ny+1, 0:nz+1) is stored x fastest; Fortran f(i,j,k) is
 * C f[i + (nx+2)*(j + (ny+2)*k)].  1-D z arrays (dzf, dzf_inv, dzc_inv,
 * dzflzi) hold indices 0..nz+1.
 *
 * Parity pins (tests/test_oracle_*.py): halfcomplex FFT vs numpy.fft.rfft;
 * eigenvalues and z coefficients vs the flux-form operator; PPE vs a dense
 * least-squares solve of the assembled operator on tiny grids and vs
 * manufactured solutions; stencils vs closed forms (1-D viscous profile,
 * constant fields, discrete momentum conservation); projection divergence.
 */
#ifndef PLL_ORACLE_H
#define PLL_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;
    double dx_inv, dy_inv;   /* 1/dx, 1/dy                          */
    double re_inv, dt;       /* viscosity nu = 1/Re, time step      */
    double lz;
    const double *dzf;       /* [nz+2] cell heights, ghosts at 0, nz+1 */
    const double *dzf_inv;   /* [nz+2]                              */
    const double *dzc_inv;   /* [nz+2] (index nz+1 unused)          */
    const double *dzflzi;    /* [nz+2] dzf/lz                       */
} orc_grid_t;

/* ---- halfcomplex FFT (FFTW_R2HC / FFTW_HC2R semantics, unnormalised) ---- */
void orc_r2hc(int n, const double *in, double *out);
void orc_hc2r(int n, const double *in, double *out);

/* ---- mesh ---- */
void orc_mesh_tanh(int nz, double lz, double stretch_ratio, double *dzf_global);
void orc_mesh_derived(int nz, double lz, const double *dzf, double *dzf_inv,
                      double *dzc, double *dzc_inv, double *dzflzi);

/* ---- PPE pieces ---- */
void orc_eigen_values(int ist, int isz, int isz_global, const char *bctype,
                      double *lambda);
void orc_set_trid_coeff(int n, const double *dzf, const char *bctype, char c_or_f,
                        const int neighbor[2], double *a, double *b, double *c);

typedef struct orc_ppe orc_ppe_t;
/* pin_gauge = 0: literal reference arithmetic (b == 0 -> 0, else divide);
 * pin_gauge = 1: the singular (0,0) Neumann mode's top value is set to 0. */
orc_ppe_t *orc_ppe_init(int nx, int ny, int nz, double dx, double dy,
                        const double *dzf_global, const char *bctype_x,
                        const char *bctype_y, const char *bctype_z, int pin_gauge);
void orc_ppe_solve(orc_ppe_t *h, double *var);
void orc_ppe_free(orc_ppe_t *h);
/* Apply the discrete Poisson operator L_h to the interior of a haloed p whose
 * ghosts satisfy the solver's boundary conditions (periodic x,y; for z the
 * ghost rows are ignored and the Neumann/Dirichlet closure of
 * set_trid_coeff is used).  Output: interior-only dense nx*ny*nz array. */
void orc_ppe_apply(orc_ppe_t *h, const double *var, double *out);
/* Seconds spent in each of the five solver stages on the last call. */
void orc_ppe_stage_times(const orc_ppe_t *h, double t[5]);

/* ---- step ---- */
void orc_transform2crf(const orc_grid_t *g, double vel_crf, double *vel);
void orc_rk1(const orc_grid_t *g, const double *u, const double *v, const double *w,
             double *u1, double *v1, double *w1, double u_crf);
void orc_rk2(const orc_grid_t *g, double *u, double *v, double *w,
             const double *u1, const double *v1, const double *w1, double u_crf);
void orc_calc_rhs(const orc_grid_t *g, const double *u, const double *v,
                  const double *w, double *rhs);
void orc_correct_vel(const orc_grid_t *g, const double *p, double *u, double *v,
                     double *w);
double orc_mean_h(const orc_grid_t *g, const double *var);
void orc_force_vel(const orc_grid_t *g, const int is_forced[3],
                   const double vel_force[3], double *u, double *v, double *w);
void orc_bc_vel(const orc_grid_t *g, double *u, double *v, double *w, double u_crf);
void orc_bc_p(const orc_grid_t *g, double *p);
double orc_max_div(const orc_grid_t *g, const double *u, const double *v,
                   const double *w, int ijk[3]);
void orc_wall_tau(const orc_grid_t *g, const double *u, double u_crf, double tau[2]);
double orc_wall_shear(const orc_grid_t *g, const double *u, double u_crf);
double orc_max_cfl(const orc_grid_t *g, const double *u, const double *v,
                   const double *w, int ijk[3]);
int orc_has_nan(long n, const double *q);
/* orc_step with a body force (bu, bv, bw) added to the momentum right-hand
 * side of both RK stages (for the adjoint checks; not in the reference) */
void orc_step_body(const orc_grid_t *g, orc_ppe_t *ppe, double *u, double *v, double *w,
                   double *u1, double *v1, double *w1, double *p, double u_crf,
                   const int is_forced[3], const double vel_force[3], const double *bu,
                   const double *bv, const double *bw);
/* One full time step in the order of src/main.f90:137-208. */
/* orc_step timed per GPTL bucket of src/main.f90:131-212 (seconds): uvw1,
 * Update boundary vel, uvw2, Calculate RHS, Poisson solver, Update boundary
 * pres, Correct vel, Force vel. */
void orc_step_timed(const orc_grid_t *g, orc_ppe_t *ppe, double *u, double *v, double *w,
                    double *u1, double *v1, double *w1, double *p, double u_crf,
                    const int is_forced[3], const double vel_force[3], double t[8]);
void orc_step(const orc_grid_t *g, orc_ppe_t *ppe, double *u, double *v, double *w,
              double *u1, double *v1, double *w1, double *p, double u_crf,
              const int is_forced[3], const double vel_force[3]);

/* threads used by the OpenMP loops (0 = library default) */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
