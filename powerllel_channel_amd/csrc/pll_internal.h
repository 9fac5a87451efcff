/*
 * Internal interface between the C host layer (ppe.c, channel.c) and the HIP
 * kernel launchers (*.hip).  Plain C structs; device pointers as void*.
 */
#ifndef PLL_INTERNAL_H
#define PLL_INTERNAL_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLL_MAX_STAGES 12

/* One complex FFT of length L done by T threads holding E points each
 * (L = T*E), in nst radix stages rad[0..nst-1]; a radix that does not divide
 * E runs its stage through LDS. */
typedef struct {
    int L, E, T, nst;
    int rad[PLL_MAX_STAGES];
} pll_fftplan;

/* Everything the PPE kernels need, all pointers on the device.
 *
 * Spectral work array W (N doubles = L*ny*nz complex, L = nx/2), layout
 * W[k][my][mx], mx fastest.  After the x transform, slot mx of a row holds
 * X_mx for 1 <= mx < L and (X_0, X_L) packed in slot 0.  After the y
 * transform, columns mx >= 1 hold the complex DFT along y; column 0 holds
 * A = DFT_y(X_0) and B = DFT_y(X_L) (both Hermitian) packed as
 *   slot 0: (A_0, B_0), slot ny/2: (A_{ny/2}, B_{ny/2}),
 *   slot my in (0, ny/2): A_my,  slot ny-my: B_my.
 * Every real number of W is then one independent tridiagonal system in z
 * with eigenvalue lambda_x(ix) + lambda_y(iy) (ix in [0,L], iy in [0,ny/2]). */
typedef struct {
    int nx, ny, nz, L;
    int pin_gauge;             /* 1: (0,0) mode top value fixed at 0 */
    double normfactor;         /* 1/(nx*ny) */
    pll_fftplan px, py;
    int x_rows_per_wg;         /* rows per workgroup, x kernels     */
    int y_cols_per_wg;         /* columns per workgroup, y kernels  */
    int y_cols32_per_wg;       /* the same for the FP32 y kernels   */
    int x_xcd;                 /* x kernels numbered per XCD: 0 none, 1 FP32, 2 all */
    int x_slab;                /* ... in y slabs per XCD: 0 none, 1 stencil sources, 2 all */
    const void *tw_x;          /* double2[nx]: exp(-2 pi i t / nx)  */
    const void *tw_y;          /* double2[ny]: exp(-2 pi i t / ny)  */
    const void *tw_x32;        /* float2[nx], same table in FP32 (refined form) */
    const void *tw_y32;        /* float2[ny] */
    double idx2, idy2;         /* 1/dx^2, 1/dy^2 (residual of the refined form) */
    const double *a, *c;       /* [nz] tridiagonal off-diagonals    */
    const double *bz;          /* [nz] diagonal before adding the eigenvalue */
    const double *dz;          /* [nz] row sums a + bz + c (zero up to rounding) */
    const double *piv;         /* [nz][ny/2+1][L+1] factorised diagonal */
    struct {                   /* partitioned z solve (P = 0: streaming solve) */
        int M, P, S;           /* rows per lane, lanes per column, columns per tile */
        int dpp16;             /* P == 16: interface reduction by DPP row shifts */
        int qx, iyx;           /* tiles with q0 < qx and iy < iyx: exact path */
        const double *bz;      /* [nz] diagonal before adding the eigenvalue */
        const double *lamx;    /* [L+1]    */
        const double *lamy;    /* [ny/2+1] */
        const int *order;      /* [tiles] launch order, exact tiles first */
        int h32;               /* FP32 storage: 2 = tiles of 2 S columns (0/1: S) */
        int h64;               /* FP64 storage: the same, at S = 16 only */
        int qx32;              /* qx for those tiles */
        const int *order32;    /* [nx/(2S) ny] their launch order */
    } tp;
    /* refined form: the lowest modes (kx in {0,1}, ky in {0,1,ny-1}) of each
     * FP32 solve in FP64 (low_on = 0: off).  low_rx [ny nz][2] the x sums of
     * the FP64 rows, low_y [nz][6] their y sums, low_x [nz][6] the solved
     * modes (double2 each) */
    int low_on;
    void *low_rx, *low_y, *low_x;
} pll_ppe_dev;

int pll_launch_xfwd(const pll_ppe_dev *d, const double *p, void *W, hipStream_t s);

/* Velocity source of the x forward transform: the PPE right-hand side of
 * calcRHS formed in the transform's load (haloed u, v, w; dzf_inv with the
 * ghost entries, indexed by the haloed k). */
typedef struct {
    const double *u, *v, *w, *dzf_inv;
    double dti, dtidxi, dtidyi;
    /* 1: the velocity's boundary fill has not been made; the transform reads
     * what it would have left (u(0) = u(nx), v(j=0) = v(ny), w = 0 on the
     * wall rows k = 0 and nz) */
    int unfilled;
} pll_div_src;
int pll_launch_xfwd_div(const pll_ppe_dev *d, const pll_div_src *q, void *W, hipStream_t s);

/* FP32 solve of the refined form.  W32 is the spectral array in FP32 (N
 * floats, same layout as W).  x forward sources and x backward destinations: */
enum { PLL_SRC_HALO64 = 0, PLL_SRC_DENSE32 = 1 };
enum {
    PLL_DST_HALO_SET = 0,        /* haloed p interior := c               */
    PLL_DST_DENSE_SET = 1,       /* dense P := c                          */
    PLL_DST_DENSE_ADD = 2,       /* dense P += c                          */
    PLL_DST_HALO_FROM_DENSE = 3  /* haloed p interior := dense P + c      */
};
int pll_launch_xfwd32(const pll_ppe_dev *d, int src_kind, const void *src, void *W32,
                      hipStream_t s);
int pll_launch_yfwd32(const pll_ppe_dev *d, void *W32, hipStream_t s);
int pll_launch_trid32(const pll_ppe_dev *d, void *W32, hipStream_t s);
int pll_launch_ybwd32(const pll_ppe_dev *d, void *W32, hipStream_t s);
/* eps: device scalar subtracted by PLL_DST_HALO_FROM_DENSE (NULL: none). */
int pll_launch_xbwd32(const pll_ppe_dev *d, int dst_kind, const void *W32, double *p_halo,
                      double *P_dense, const double *eps, hipStream_t s);
/* r = f - L_h P in FP64, stored as FP32 (dense, x fastest); f is the interior
 * of the haloed array, P dense.  gauge non-NULL (pll_gauge_parts(d) + 1
 * doubles): also gauge[parts] := mean of P over the plane k = nz-1. */
/* Refinement with a tolerance (the solution lives in the caller's haloed
 * array, the right-hand side is kept in a dense FP64 copy F):
 * x forward of the haloed f that also writes F; x backward into the haloed
 * p, p := (p - *eps) + c, reporting in norm[0] and norm[1] max |c| and
 * max |p| (norm holds 2 + 2*ny*nz doubles). */
int pll_launch_xfwd32_copy(const pll_ppe_dev *d, const double *f_halo, double *F, void *W32,
                           hipStream_t s);
int pll_launch_xbwd32_norm(const pll_ppe_dev *d, const void *W32, double *p_halo,
                           const double *eps, double *norm, hipStream_t s);
/* x forward (FP32) of the residual r = f - L_h p formed in the load (f and p
 * haloed when the flag is set, dense otherwise); same values as
 * pll_launch_resid followed by the dense FP32 x forward */
int pll_launch_xfwd32_resid(const pll_ppe_dev *d, const double *f, int f_halo, const double *P,
                            int p_halo, void *W32, hipStream_t s);
/* gauge[parts] := top-plane mean of p (haloed or dense), from the same
 * per-block partial sums as pll_launch_resid computes */
int pll_launch_gauge(const pll_ppe_dev *d, const double *P, int halo, double *gauge,
                     hipStream_t s);
/* residual with a dense f and a haloed p (see pll_launch_resid) */
int pll_launch_resid_hp(const pll_ppe_dev *d, const double *F, const double *p_halo, float *r,
                        double *gauge, hipStream_t s);
int pll_launch_resid(const pll_ppe_dev *d, const double *f_halo, const double *P, float *r,
                     double *gauge, hipStream_t s);
int pll_gauge_parts(const pll_ppe_dev *d);
/* Lowest modes of the refined form in FP64, after an x forward that wrote
 * low_rx: their y sums, their tridiagonal solves (the reference's recurrence
 * with the stored pivots and the pinned gauge) into low_x; and the same
 * modes set to zero in the FP32 spectrum W32 (after its y forward). */
int pll_launch_low_solve(const pll_ppe_dev *d, hipStream_t s);
int pll_launch_low_zero(const pll_ppe_dev *d, void *W32, hipStream_t s);
int pll_launch_yfwd(const pll_ppe_dev *d, void *W, hipStream_t s);
int pll_launch_trid(const pll_ppe_dev *d, void *W, hipStream_t s);
int pll_launch_ybwd(const pll_ppe_dev *d, void *W, hipStream_t s);
int pll_launch_xbwd(const pll_ppe_dev *d, const void *W, double *p, hipStream_t s);
/* x transforms of the rows [row0, row1) only (row = k ny + j): the
 * transfer pipeline of host-pointer solves runs them per chunk of planes. */
int pll_launch_xfwd_rows(const pll_ppe_dev *d, const double *p, void *W, long long row0,
                         long long row1, hipStream_t s);
int pll_launch_xbwd_rows(const pll_ppe_dev *d, const void *W, double *p, long long row0,
                         long long row1, hipStream_t s);
/* host memcpy split over a pool of threads (host-pointer solves) */
typedef struct pll_copy_pool pll_copy_pool;
int pll_copy_pool_create(pll_copy_pool **out, int nthreads);
void pll_copy_pool_memcpy(pll_copy_pool *p, void *dst, const void *src, size_t n);
void pll_copy_pool_destroy(pll_copy_pool *p);
/* pll_launch_xbwd that also writes p's boundary fill (updateBoundP). */
int pll_launch_xbwd_fill(const pll_ppe_dev *d, const void *W, double *p, hipStream_t s);
/* 1 if the partitioned z solve is instantiated for M rows per lane. */
int pll_trid_part_supported(int M);
/* 1 if a kernel instantiation exists for E (elements per thread). */
int pll_fft_supported_E(int E);
/* largest workgroup the E instantiation is compiled for (FP64 kernels;
 * the FP32 ones take pll_fft_max_threads32) */
int pll_fft_max_threads(int E);
int pll_fft_max_threads32(int E);
/* compile-time plan for length L, if one exists (returns 1 and fills p);
 * among several, the one with E == prefer_E, else the first listed */
int pll_fft_ct_plan(int L, int prefer_E, pll_fftplan *p);
/* the plan the solver uses for length L (host only): E, stage count, radices */
int pll_fft_plan_info(int L, int *E, int *nst, int *rad);

/* Grid constants for the step kernels (device pointers for z arrays). */
typedef struct {
    int nx, ny, nz;
    double dx_inv, dy_inv, re_inv, dt, u_crf;
    const double *dzf, *dzf_inv, *dzc_inv, *dzflzi; /* [nz+2] */
    int rk_kc;                 /* planes per k chunk of the stencil tiles (0: 64) */
} pll_grid_dev;

/* body: NULL, or three device fields (haloed layout) added to the momentum
 * right-hand side (a body force; not part of the reference's step). */
int pll_launch_rk1(const pll_grid_dev *g, const double *u, const double *v,
                   const double *w, double *u1, double *v1, double *w1,
                   const double *const *body, hipStream_t s);
/* RK1 from an unprojected velocity (u, v, w: the previous RK2 output after its
 * boundary fill; p: its pressure with ghosts): every value read is projected
 * on load as correctVel + forceVel (fmask bits with vfu, vfv and the device
 * means mean[0..1]; mean NULL: no forcing) + updateBoundVel would leave it;
 * the projected interior is written to (nu, nv, nw), the RK1 result to
 * (u1, v1, w1). */
int pll_launch_rk1_corr(const pll_grid_dev *g, const double *u, const double *v,
                        const double *w, const double *p, double *nu, double *nv, double *nw,
                        double *u1, double *v1, double *w1, int fmask, double vfu, double vfv,
                        const double *mean, hipStream_t s);
/* mpart non-NULL and fmask != 0: also the Mean_h partial sums of the new u
 * (fmask bit 0) and v (bit 1), pll_rk_blocks(g) of each, u's first. */
/* fill != 0: (u1, v1, w1) have no boundary fill; RK2 reads the values the
 * fill (updateBoundVel) would leave from their interior. */
int pll_launch_rk2(const pll_grid_dev *g, double *u, double *v, double *w,
                   const double *u1, const double *v1, const double *w1, double *mpart,
                   int fmask, const double *const *body, int fill, hipStream_t s);
/* RK2 of a multi-step run in place on (u, v, w), the previous step's
 * unprojected RK2 output, projected with p (and the forcing shift from mean)
 * as it is read; Mean_h partials into mpart as pll_launch_rk2 */
int pll_launch_rk2_proj(const pll_grid_dev *g, double *u, double *v, double *w,
                        const double *u1, const double *v1, const double *w1,
                        const double *p, double *mpart, int fmask, double vfu,
                        double vfv, const double *mean, hipStream_t s);
int pll_rk_blocks(const pll_grid_dev *g);
int pll_launch_calc_rhs(const pll_grid_dev *g, const double *u, const double *v,
                        const double *w, double *p, hipStream_t s);
/* mean non-NULL and fmask != 0: then also q += vel_force - mean[d] for the
 * forced components (bit 0 u with vfu, bit 1 v with vfv). */
int pll_launch_correct_vel(const pll_grid_dev *g, const double *p, double *u,
                           double *v, double *w, int fmask, double vfu, double vfv,
                           const double *mean, hipStream_t s);
/* partial sums of q*dzflzi per (j,k) row into part[ny*nz]; then a single
 * workgroup reduces them to out[0] = sum / nx / ny (Mean_h). */
int pll_launch_mean_partials(const pll_grid_dev *g, const double *q, double *part,
                             hipStream_t s);
int pll_launch_mean_final(int n, const double *part, double *out, int nx, int ny,
                          hipStream_t s);
/* q += (vel_force - mean) on the interior, mean read from device *mean. */
int pll_launch_force_shift(const pll_grid_dev *g, double *q, double vel_force,
                           const double *mean, hipStream_t s);
int pll_launch_bc_vel(const pll_grid_dev *g, double *u, double *v, double *w,
                      hipStream_t s);
int pll_launch_bc_p(const pll_grid_dev *g, double *p, hipStream_t s);
/* per-(j,k)-row max |div| with linear index; then final reduction. */
int pll_launch_maxdiv_partials(const pll_grid_dev *g, const double *u,
                               const double *v, const double *w, double *pval,
                               long long *pidx, hipStream_t s);
int pll_launch_maxdiv_final(int n, const double *pval, const long long *pidx,
                            double *oval, long long *oidx, hipStream_t s);
/* the same reduction for the CFL sum (calcMaxCFL before the factor dt) */
int pll_launch_maxcfl_partials(const pll_grid_dev *g, const double *u,
                               const double *v, const double *w, double *pval,
                               long long *pidx, hipStream_t s);
/* flags[0 .. nblocks) = 1 where a block's share of the haloed field holds a NaN */
int pll_launch_nan_partials(const pll_grid_dev *g, const double *q, long long *flags,
                            int nblocks, hipStream_t s);

/* ---- adjoint pieces (adjoint_kernels.hip); every array is haloed ---- */
/* o = beta o + a0 l [interior] + alpha J_f(q)^T l, over all cells; l read on
 * the interior (J_f: Jacobian of the momentum right-hand side). */
int pll_launch_jtv(const pll_grid_dev *g, const double *const q[3], const double *const l[3],
                   double *const o[3], double alpha, double a0, double beta, hipStream_t s);
/* o = J_f(q) dq on the interior. */
int pll_launch_jvp(const pll_grid_dev *g, const double *const q[3], const double *const dq[3],
                   double *const o[3], hipStream_t s);
/* transposes of the boundary fills: ghost adjoints folded into the interior
 * values they were made from, ghosts cleared */
int pll_launch_bc_vel_adj(const pll_grid_dev *g, double *lu, double *lv, double *lw,
                          hipStream_t s);
int pll_launch_bc_p_adj(const pll_grid_dev *g, double *lp, hipStream_t s);
/* lp := transpose of correctVel's pressure term applied to (lu, lv, lw) */
int pll_launch_correct_adj(const pll_grid_dev *g, const double *lu, const double *lv,
                           const double *lw, double *lp, hipStream_t s);
/* (lu, lv, lw) += transpose of calcRHS applied to lr (interior) */
int pll_launch_rhs_adj(const pll_grid_dev *g, const double *lr, double *lu, double *lv,
                       double *lw, hipStream_t s);
/* q -= dzflzi(k) sum_interior(q) / (nx ny); part: ny*nz, total: 1 double */
int pll_launch_force_adj(const pll_grid_dev *g, double *q, double *part, double *total,
                         hipStream_t s);
/* out[0], out[1] = sum over (i,j) of u at k = 1 and k = nz; part: 2 ny */
int pll_launch_wall_sum(const pll_grid_dev *g, const double *u, double *part, double *out,
                        hipStream_t s);
/* lu += c_bot on the plane k = 1 and c_top on k = nz (interior) */
int pll_launch_wall_seed(const pll_grid_dev *g, double *lu, double c_bot, double c_top,
                         hipStream_t s);
/* y += a x over every cell, three fields */
int pll_launch_axpy3(const pll_grid_dev *g, double a, const double *const x[3],
                     double *const y[3], hipStream_t s);

/* Host-layer calls between ppe.c and channel.c (pll_ppe from powerllel.h). */
struct pll_ppe;
/* FP64 solve whose right-hand side is formed from the velocity inside the x
 * forward transform (calcRHS fused into the solve); the solution lands in the
 * interior of d_var and, unlike pll_ppe_solve, its boundary fill
 * (updateBoundP) is written too.  Only for the FP64 form. */
int pll_ppe_solve_div(struct pll_ppe *h, const pll_div_src *q, double *d_var, void *stream);
/* sweeps of the refined form that pll_ppe_solve uses (-1: FP64 solve). */
int pll_ppe_refinement(const struct pll_ppe *h);
struct pll_channel;
const pll_grid_dev *pll_channel_grid(const struct pll_channel *c);
int pll_channel_forcing(const struct pll_channel *c, int is_forced[3], double vel_force[3]);
const double *pll_channel_dzf_host(const struct pll_channel *c);
/* averaged wall shear from the two plane sums of u */
double pll_wall_shear_from_sums(const struct pll_channel *c, const double sums[2]);
int pll_channel_step_body_internal(struct pll_channel *c, struct pll_ppe *ppe, double *u,
                                   double *v, double *w, double *u1, double *v1, double *w1,
                                   double *p, const double *const *body, void *stream);

#ifdef __cplusplus
}
#endif
#endif
