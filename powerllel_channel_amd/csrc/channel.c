/*
 * Host side of the RK2 projection step on the GPU (C, drives the HIP kernels).
 *
 * One handle holds the grid (mod_mesh.f90:26-84 derived arrays on the
 * device), the physics constants and the small reduction buffers.  The
 * entry points mirror the reference's per-step routines (src/main.f90:131-212).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/powerllel.h"
#include "pll_internal.h"

int pll_set_error(int code, const char *fmt, ...);

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return pll_set_error(PLL_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct pll_channel {
    pll_grid_dev g;
    int is_forced[3];
    double vel_force[3];
    double *zarr;      /* device: dzf | dzf_inv | dzc_inv | dzflzi, each nz+2 */
    double *part;      /* device: ny*nz partials */
    long long *pidx;   /* device: ny*nz indices  */
    double *scal;      /* device: 8 scalars (mean, max value, fused-step means at 4, 5) */
    long long *sidx;   /* device: 1 index */
    double *dzf_h;     /* host copy of dzf_global (nz+2) */
    double lx, ly, lz;
};

static int kl(int rc, const char *what)
{
    if (rc == 0) return PLL_OK;
    return pll_set_error(PLL_EHIP, "%s: kernel launch failed: %s", what,
                         hipGetErrorString(hipGetLastError()));
}

int pll_channel_destroy(pll_channel *c)
{
    if (!c) return PLL_OK;
    if (c->zarr) hipFree(c->zarr);
    if (c->part) hipFree(c->part);
    if (c->pidx) hipFree(c->pidx);
    if (c->scal) hipFree(c->scal);
    if (c->sidx) hipFree(c->sidx);
    free(c->dzf_h);
    free(c);
    return PLL_OK;
}

int pll_channel_create(pll_channel **out, int nx, int ny, int nz, double lx, double ly,
                       double lz, const double *dzf_global, double re_inv, double dt,
                       double u_crf, const int is_forced[3], const double vel_force[3])
{
    if (!out) return pll_set_error(PLL_EINVAL, "out is NULL");
    *out = NULL;
    if (nx < 2 || ny < 2 || nz < 2 || !dzf_global)
        return pll_set_error(PLL_EINVAL, "bad grid (%d %d %d)", nx, ny, nz);
    if (!(lx > 0) || !(ly > 0) || !(lz > 0) || !(dt > 0))
        return pll_set_error(PLL_EINVAL, "lengths and dt must be > 0");
    for (int k = 0; k < nz + 2; k++)
        if (!(dzf_global[k] > 0.0)) return pll_set_error(PLL_EINVAL, "dzf_global[%d] <= 0", k);
    pll_channel *c = (pll_channel *)calloc(1, sizeof *c);
    if (!c) return pll_set_error(PLL_ENOMEM, "calloc failed");
    const int n2 = nz + 2;
    double *z = (double *)calloc(4 * (size_t)n2, sizeof(double));
    if (!z) {
        free(c);
        return pll_set_error(PLL_ENOMEM, "calloc failed");
    }
    double *dzf = z, *dzf_inv = z + n2, *dzc_inv = z + 2 * n2, *dzflzi = z + 3 * n2;
    /* mod_mesh.f90:42-45, 64-72 (one rank: dzf = dzf_global) */
    for (int k = 0; k < n2; k++) {
        dzf[k] = dzf_global[k];
        dzf_inv[k] = 1.0 / dzf[k];
    }
    for (int k = 0; k <= nz; k++) {
        double dzc = 0.5 * (dzf[k] + dzf[k + 1]);
        dzc_inv[k] = 1.0 / dzc;
    }
    for (int k = 0; k < n2; k++) dzflzi[k] = dzf[k] / lz;
    const double dx = lx / nx, dy = ly / ny;
    c->g.nx = nx;
    c->g.ny = ny;
    c->g.nz = nz;
    c->g.dx_inv = 1.0 / dx;
    c->g.dy_inv = 1.0 / dy;
    c->g.re_inv = re_inv;
    c->g.dt = dt;
    c->g.u_crf = u_crf;
    {
        /* planes per k chunk of the stencil kernels' tiles (tuning override;
         * fixed for the handle's life: the Mean_h partials are sized by it) */
        const char *ek = getenv("PLL_RK_KC");
        c->g.rk_kc = ek && atoi(ek) > 0 ? atoi(ek) : 0;
    }
    for (int d = 0; d < 3; d++) {
        c->is_forced[d] = is_forced ? is_forced[d] != 0 : (d == 0);
        c->vel_force[d] = vel_force ? vel_force[d] : (d == 0 ? 1.0 : 0.0);
    }
    hipError_t e;
    size_t nrow = (size_t)ny * nz;
    const size_t nfused = 2 * (size_t)pll_rk_blocks(&c->g); /* fused-step Mean_h partials */
    if (nfused > nrow) nrow = nfused;
    if ((e = hipMalloc((void **)&c->zarr, sizeof(double) * 4 * (size_t)n2)) != hipSuccess ||
        (e = hipMalloc((void **)&c->part, sizeof(double) * nrow)) != hipSuccess ||
        (e = hipMalloc((void **)&c->pidx, sizeof(long long) * nrow)) != hipSuccess ||
        (e = hipMalloc((void **)&c->scal, sizeof(double) * 8)) != hipSuccess ||
        (e = hipMalloc((void **)&c->sidx, sizeof(long long))) != hipSuccess ||
        (e = hipMemcpy(c->zarr, z, sizeof(double) * 4 * (size_t)n2, hipMemcpyHostToDevice)) !=
            hipSuccess) {
        free(z);
        pll_channel_destroy(c);
        return pll_set_error(PLL_ENOMEM, "device allocation/copy failed: %s", hipGetErrorString(e));
    }
    free(z);
    c->dzf_h = (double *)malloc(sizeof(double) * (size_t)n2);
    if (!c->dzf_h) {
        pll_channel_destroy(c);
        return pll_set_error(PLL_ENOMEM, "malloc failed");
    }
    memcpy(c->dzf_h, dzf_global, sizeof(double) * (size_t)n2);
    c->lx = lx;
    c->ly = ly;
    c->lz = lz;
    c->g.dzf = c->zarr;
    c->g.dzf_inv = c->zarr + n2;
    c->g.dzc_inv = c->zarr + 2 * n2;
    c->g.dzflzi = c->zarr + 3 * n2;
    *out = c;
    return PLL_OK;
}

#define CHK_C(c) \
    if (!(c)) return pll_set_error(PLL_ESTATE, "channel handle is NULL")

int pll_channel_rk1(pll_channel *c, const double *u, const double *v, const double *w,
                    double *u1, double *v1, double *w1, void *stream)
{
    CHK_C(c);
    return kl(pll_launch_rk1(&c->g, u, v, w, u1, v1, w1, NULL, (hipStream_t)stream), "rk1");
}

int pll_channel_rk2(pll_channel *c, double *u, double *v, double *w, const double *u1,
                    const double *v1, const double *w1, void *stream)
{
    CHK_C(c);
    return kl(pll_launch_rk2(&c->g, u, v, w, u1, v1, w1, NULL, 0, NULL, 0, (hipStream_t)stream),
              "rk2");
}

int pll_channel_calc_rhs(pll_channel *c, const double *u, const double *v, const double *w,
                         double *p, void *stream)
{
    CHK_C(c);
    return kl(pll_launch_calc_rhs(&c->g, u, v, w, p, (hipStream_t)stream), "calc_rhs");
}

int pll_channel_correct_vel(pll_channel *c, const double *p, double *u, double *v, double *w,
                            void *stream)
{
    CHK_C(c);
    return kl(pll_launch_correct_vel(&c->g, p, u, v, w, 0, 0.0, 0.0, NULL, (hipStream_t)stream),
              "correct_vel");
}

static int mean_dev(pll_channel *c, const double *q, hipStream_t s)
{
    int rc;
    if ((rc = kl(pll_launch_mean_partials(&c->g, q, c->part, s), "mean partials"))) return rc;
    return kl(pll_launch_mean_final(c->g.ny * c->g.nz, c->part, c->scal, c->g.nx, c->g.ny, s),
              "mean final");
}

int pll_channel_force_vel(pll_channel *c, double *u, double *v, double *w, void *stream)
{
    CHK_C(c);
    hipStream_t s = (hipStream_t)stream;
    double *q3[3] = {u, v, w};
    for (int d = 0; d < 3; d++) {
        if (!c->is_forced[d]) continue;
        int rc;
        if ((rc = mean_dev(c, q3[d], s))) return rc;
        if ((rc = kl(pll_launch_force_shift(&c->g, q3[d], c->vel_force[d], c->scal, s),
                     "force shift")))
            return rc;
    }
    return PLL_OK;
}

int pll_channel_bc_vel(pll_channel *c, double *u, double *v, double *w, void *stream)
{
    CHK_C(c);
    return kl(pll_launch_bc_vel(&c->g, u, v, w, (hipStream_t)stream), "bc_vel");
}

int pll_channel_bc_p(pll_channel *c, double *p, void *stream)
{
    CHK_C(c);
    return kl(pll_launch_bc_p(&c->g, p, (hipStream_t)stream), "bc_p");
}

int pll_channel_mean_h(pll_channel *c, const double *q, double *mean, void *stream)
{
    CHK_C(c);
    if (!mean) return pll_set_error(PLL_EINVAL, "mean is NULL");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = mean_dev(c, q, s))) return rc;
    HIPCHK(hipMemcpyAsync(mean, c->scal, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return PLL_OK;
}

int pll_channel_max_div(pll_channel *c, const double *u, const double *v, const double *w,
                        double *value, int ijk[3], void *stream)
{
    CHK_C(c);
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = kl(pll_launch_maxdiv_partials(&c->g, u, v, w, c->part, c->pidx, s), "maxdiv")))
        return rc;
    if ((rc = kl(pll_launch_maxdiv_final(c->g.ny * c->g.nz, c->part, c->pidx, c->scal + 1,
                                         c->sidx, s),
                 "maxdiv final")))
        return rc;
    double val;
    long long idx;
    HIPCHK(hipMemcpyAsync(&val, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&idx, c->sidx, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (value) *value = val < 0.0 ? 0.0 : val;
    if (ijk) {
        const long long nx = c->g.nx, ny = c->g.ny;
        ijk[0] = (int)(idx % nx) + 1;
        ijk[1] = (int)((idx / nx) % ny) + 1;
        ijk[2] = (int)(idx / (nx * ny)) + 1;
    }
    return PLL_OK;
}

/* calcMaxCFL, mod_utils.f90:193-255: max over the interior of the CFL sum
 * times dt, and its 1-based (i,j,k). */
int pll_channel_max_cfl(pll_channel *c, const double *u, const double *v, const double *w,
                        double *value, int ijk[3], void *stream)
{
    CHK_C(c);
    if (!u || !v || !w) return pll_set_error(PLL_EINVAL, "NULL field");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = kl(pll_launch_maxcfl_partials(&c->g, u, v, w, c->part, c->pidx, s), "maxcfl")))
        return rc;
    if ((rc = kl(pll_launch_maxdiv_final(c->g.ny * c->g.nz, c->part, c->pidx, c->scal + 1,
                                         c->sidx, s),
                 "maxcfl final")))
        return rc;
    double val;
    long long idx;
    HIPCHK(hipMemcpyAsync(&val, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&idx, c->sidx, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (value) *value = (val < 0.0 ? 0.0 : val) * c->g.dt;
    if (ijk) {
        const long long nx = c->g.nx, ny = c->g.ny;
        ijk[0] = (int)(idx % nx) + 1;
        ijk[1] = (int)((idx / nx) % ny) + 1;
        ijk[2] = (int)(idx / (nx * ny)) + 1;
    }
    return PLL_OK;
}

/* checkNaN, mod_utils.f90:352-394, over every element of the haloed field. */
int pll_channel_check_nan(pll_channel *c, const double *q, int *has_nan, void *stream)
{
    CHK_C(c);
    if (!q || !has_nan) return pll_set_error(PLL_EINVAL, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    int nb = c->g.ny * c->g.nz;
    if (nb > 1024) nb = 1024;
    int rc;
    if ((rc = kl(pll_launch_nan_partials(&c->g, q, c->pidx, nb, s), "checkNaN"))) return rc;
    long long flags[1024];
    HIPCHK(hipMemcpyAsync(flags, c->pidx, sizeof(long long) * (size_t)nb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int any = 0;
    for (int b = 0; b < nb; b++) any |= flags[b] != 0;
    *has_nan = any;
    return PLL_OK;
}

static int step_impl(pll_channel *c, pll_ppe *ppe, double *u, double *v, double *w, double *u1,
                     double *v1, double *w1, double *p, const double *const *body, void *stream)
{
    CHK_C(c);
    if (!ppe) return pll_set_error(PLL_ESTATE, "PPE handle is NULL");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    /* Forcing of u and/or v is fused: Mean_h is reduced inside RK2 and the
     * shift applied inside correctVel.  correctVel changes u by
     * -(p_{i+1} - p_i) dt/dx, whose sum over a periodic x row telescopes to
     * zero (and likewise v in y), so Mean_h of the RK2 output equals Mean_h
     * of the corrected field in exact arithmetic (they differ by roundoff in
     * the summation).  A forced w keeps the reference's separate passes. */
    const int fmask = (c->is_forced[0] ? 1 : 0) | (c->is_forced[1] ? 2 : 0);
    const int fuse = fmask && !c->is_forced[2];
    const int none = !fmask && !c->is_forced[2];
    const int nb = pll_rk_blocks(&c->g);
    if ((rc = kl(pll_launch_rk1(&c->g, u, v, w, u1, v1, w1, body, s), "rk1"))) return rc;
    /* RK2 reads u1's boundary values from its interior (no fill pass) */
    if ((rc = kl(pll_launch_rk2(&c->g, u, v, w, u1, v1, w1, fuse ? c->part : NULL,
                                fuse ? fmask : 0, body, 1, s),
                 "rk2")))
        return rc;
    if (fuse)
        for (int d = 0; d < 2; d++)
            if ((fmask >> d) & 1)
                if ((rc = kl(pll_launch_mean_final(nb, c->part + (size_t)d * nb, c->scal + 4 + d,
                                                   c->g.nx, c->g.ny, s),
                             "mean final")))
                    return rc;
    if ((rc = pll_channel_bc_vel(c, u, v, w, stream))) return rc;
    if (pll_ppe_refinement(ppe) < 0) {
        /* calcRHS fused into the solve's x forward transform: the right-hand
         * side is never stored (mod_calcRHS.f90:28-30, same arithmetic) */
        const double dti = 1.0 / c->g.dt;
        pll_div_src q = {u, v, w, c->g.dzf_inv, dti, dti * c->g.dx_inv, dti * c->g.dy_inv};
        if ((rc = pll_ppe_solve_div(ppe, &q, p, stream))) return rc;  /* + updateBoundP */
    } else {
        if ((rc = pll_channel_calc_rhs(c, u, v, w, p, stream))) return rc;
        if ((rc = pll_ppe_solve(ppe, p, stream))) return rc;
        if ((rc = pll_channel_bc_p(c, p, stream))) return rc;
    }
    if (fuse || none) {
        if ((rc = kl(pll_launch_correct_vel(&c->g, p, u, v, w, fuse ? fmask : 0, c->vel_force[0],
                                            c->vel_force[1], fuse ? c->scal + 4 : NULL, s),
                     "correct_vel")))
            return rc;
    } else {
        if ((rc = pll_channel_correct_vel(c, p, u, v, w, stream))) return rc;
        if ((rc = pll_channel_force_vel(c, u, v, w, stream))) return rc;
    }
    return pll_channel_bc_vel(c, u, v, w, stream);
}

int pll_channel_step(pll_channel *c, pll_ppe *ppe, double *u, double *v, double *w, double *u1,
                     double *v1, double *w1, double *p, void *stream)
{
    return step_impl(c, ppe, u, v, w, u1, v1, w1, p, NULL, stream);
}

/* A step of a chain up to its pressure solve, in place on X = (u, v, w):
 * proj == 0: X is projected (the caller's state), plain RK1 and RK2;
 * proj == 1: X is the previous step's RK2 output, unprojected (its boundary
 * fill not made either), and p that step's pressure: RK1 projects every
 * value it reads (pll_launch_rk1_corr, writing only the RK1 stage) and RK2
 * projects its old value at each cell as it reads it (pll_launch_rk2_proj),
 * both from p and the forcing means of the previous step.  Then the forcing
 * means of the new RK2 output and the FP64 solve with calcRHS formed in its
 * x transform.  The projection of X and its boundary fill are left to the
 * next step (or to the chain's final projection). */
static int chain_step(pll_channel *c, pll_ppe *ppe, int proj, double *const *X, double *u1,
                      double *v1, double *w1, double *p, int fmask, hipStream_t s)
{
    const int nb = pll_rk_blocks(&c->g);
    const double *mean = fmask ? c->scal + 4 : NULL;
    const double dti = 1.0 / c->g.dt;
    int rc;
    if (!proj) {
        if ((rc = kl(pll_launch_rk1(&c->g, X[0], X[1], X[2], u1, v1, w1, NULL, s), "rk1")))
            return rc;
        if ((rc = kl(pll_launch_rk2(&c->g, X[0], X[1], X[2], u1, v1, w1, fmask ? c->part : NULL,
                                    fmask, NULL, 1, s),
                     "rk2")))
            return rc;
    } else {
        if ((rc = kl(pll_launch_rk1_corr(&c->g, X[0], X[1], X[2], p, NULL, NULL, NULL, u1, v1,
                                         w1, fmask, c->vel_force[0], c->vel_force[1], mean, s),
                     "rk1 (projecting)")))
            return rc;
        if ((rc = kl(pll_launch_rk2_proj(&c->g, X[0], X[1], X[2], u1, v1, w1, p,
                                         fmask ? c->part : NULL, fmask, c->vel_force[0],
                                         c->vel_force[1], mean, s),
                     "rk2 (projecting)")))
            return rc;
    }
    for (int d = 0; d < 2; d++)
        if ((fmask >> d) & 1)
            if ((rc = kl(pll_launch_mean_final(nb, c->part + (size_t)d * nb, c->scal + 4 + d,
                                               c->g.nx, c->g.ny, s),
                         "mean final")))
                return rc;
    /* no boundary fill of X: the solve's calcRHS and the next step's RK1
     * read the values the fill would leave from the interior */
    pll_div_src q = {X[0], X[1], X[2], c->g.dzf_inv, dti, dti * c->g.dx_inv,
                     dti * c->g.dy_inv, 1};
    return pll_ppe_solve_div(ppe, &q, p, s);
}

/* m steps whose projections are carried into the next step's RK1 and RK2:
 * step n leaves its RK2 output and pressure p unprojected in place; the
 * next step's RK1 and RK2 read both and project every value they read.
 * After the last step the final projection (correctVel + forcing + boundary
 * fill) is made in place.  Same kernels and roundings as m calls of
 * step_impl. */
static int run_chain(pll_channel *c, pll_ppe *ppe, int m, double *u, double *v, double *w,
                     double *u1, double *v1, double *w1, double *p, int fmask, hipStream_t s)
{
    double *X[3] = {u, v, w};
    const double *mean = fmask ? c->scal + 4 : NULL;
    int rc;
    for (int i = 0; i < m; i++)
        if ((rc = chain_step(c, ppe, i > 0, X, u1, v1, w1, p, fmask, s))) return rc;
    if ((rc = kl(pll_launch_correct_vel(&c->g, p, u, v, w, fmask, c->vel_force[0],
                                        c->vel_force[1], mean, s),
                 "correct_vel")))
        return rc;
    return pll_channel_bc_vel(c, u, v, w, s);
}

int pll_channel_run(pll_channel *c, pll_ppe *ppe, int nsteps, double *u, double *v, double *w,
                    double *u1, double *v1, double *w1, double *p, void *stream)
{
    CHK_C(c);
    if (!ppe) return pll_set_error(PLL_ESTATE, "PPE handle is NULL");
    if (nsteps < 0) return pll_set_error(PLL_EINVAL, "nsteps = %d < 0", nsteps);
    if (!u || !v || !w || !u1 || !v1 || !w1 || !p) return pll_set_error(PLL_EINVAL, "NULL field");
    hipStream_t s = (hipStream_t)stream;
    const int fmask = (c->is_forced[0] ? 1 : 0) | (c->is_forced[1] ? 2 : 0);
    /* The projecting RK1 issues more loads per value than RK1 and correctVel
     * together and needs enough workgroups to hide them: at 768x384x256
     * (4608 tiles) the chain is 2% faster per step, at 192x192x128 (288
     * tiles) 10% slower, so small grids take the per-step path. */
    const char *pc = getenv("PLL_RUN_CHAIN"); /* tuning: 0 never, 1 always */
    const int big = pll_rk_blocks(&c->g) >= 2048;
    const int chain = !c->is_forced[2] && pll_ppe_refinement(ppe) < 0 &&
                      (pc ? atoi(pc) != 0 : big);
    int rc;
    if (!chain || nsteps < 3) {
        for (int i = 0; i < nsteps; i++)
            if ((rc = step_impl(c, ppe, u, v, w, u1, v1, w1, p, NULL, stream))) return rc;
        return PLL_OK;
    }
    return run_chain(c, ppe, nsteps, u, v, w, u1, v1, w1, p, fmask, s);
}

int pll_channel_step_body_internal(pll_channel *c, pll_ppe *ppe, double *u, double *v,
                                   double *w, double *u1, double *v1, double *w1, double *p,
                                   const double *const *body, void *stream)
{
    return step_impl(c, ppe, u, v, w, u1, v1, w1, p, body, stream);
}

int pll_channel_step_body(pll_channel *c, pll_ppe *ppe, double *u, double *v, double *w,
                          double *u1, double *v1, double *w1, double *p, const double *fu,
                          const double *fv, const double *fw, void *stream)
{
    if (!fu || !fv || !fw) return pll_set_error(PLL_EINVAL, "body force field is NULL");
    const double *body[3] = {fu, fv, fw};
    return step_impl(c, ppe, u, v, w, u1, v1, w1, p, body, stream);
}

/* ------------------------------------------------------------------ adjoint pieces */
const pll_grid_dev *pll_channel_grid(const pll_channel *c) { return c ? &c->g : NULL; }

int pll_channel_forcing(const pll_channel *c, int is_forced[3], double vel_force[3])
{
    CHK_C(c);
    for (int d = 0; d < 3; d++) {
        if (is_forced) is_forced[d] = c->is_forced[d];
        if (vel_force) vel_force[d] = c->vel_force[d];
    }
    return PLL_OK;
}

const double *pll_channel_dzf_host(const pll_channel *c) { return c ? c->dzf_h : NULL; }

int pll_channel_bc_vel_adj(pll_channel *c, double *lu, double *lv, double *lw, void *stream)
{
    CHK_C(c);
    if (!lu || !lv || !lw) return pll_set_error(PLL_EINVAL, "NULL field");
    return kl(pll_launch_bc_vel_adj(&c->g, lu, lv, lw, (hipStream_t)stream), "bc_vel adjoint");
}

int pll_channel_bc_p_adj(pll_channel *c, double *lp, void *stream)
{
    CHK_C(c);
    if (!lp) return pll_set_error(PLL_EINVAL, "NULL field");
    return kl(pll_launch_bc_p_adj(&c->g, lp, (hipStream_t)stream), "bc_p adjoint");
}

int pll_channel_correct_vel_adj(pll_channel *c, const double *lu, const double *lv,
                                const double *lw, double *lp, void *stream)
{
    CHK_C(c);
    if (!lu || !lv || !lw || !lp) return pll_set_error(PLL_EINVAL, "NULL field");
    return kl(pll_launch_correct_adj(&c->g, lu, lv, lw, lp, (hipStream_t)stream),
              "correct_vel adjoint");
}

int pll_channel_calc_rhs_adj(pll_channel *c, const double *lr, double *lu, double *lv, double *lw,
                             void *stream)
{
    CHK_C(c);
    if (!lr || !lu || !lv || !lw) return pll_set_error(PLL_EINVAL, "NULL field");
    return kl(pll_launch_rhs_adj(&c->g, lr, lu, lv, lw, (hipStream_t)stream), "calc_rhs adjoint");
}

int pll_channel_force_vel_adj(pll_channel *c, double *lu, double *lv, double *lw, void *stream)
{
    CHK_C(c);
    if (!lu || !lv || !lw) return pll_set_error(PLL_EINVAL, "NULL field");
    double *q3[3] = {lu, lv, lw};
    for (int d = 0; d < 3; d++) {
        if (!c->is_forced[d]) continue;
        int rc;
        if ((rc = kl(pll_launch_force_adj(&c->g, q3[d], c->part, c->scal + 6,
                                          (hipStream_t)stream),
                     "force_vel adjoint")))
            return rc;
    }
    return PLL_OK;
}

int pll_channel_rk_jvp(pll_channel *c, const double *qu, const double *qv, const double *qw,
                       const double *du, const double *dv, const double *dw, double *ou,
                       double *ov, double *ow, void *stream)
{
    CHK_C(c);
    const double *q[3] = {qu, qv, qw}, *d[3] = {du, dv, dw};
    double *o[3] = {ou, ov, ow};
    for (int i = 0; i < 3; i++)
        if (!q[i] || !d[i] || !o[i]) return pll_set_error(PLL_EINVAL, "NULL field");
    return kl(pll_launch_jvp(&c->g, q, d, o, (hipStream_t)stream), "rk jvp");
}

int pll_channel_rk_vjp(pll_channel *c, const double *qu, const double *qv, const double *qw,
                       const double *lu, const double *lv, const double *lw, double *ou,
                       double *ov, double *ow, double alpha, double a0, double beta,
                       void *stream)
{
    CHK_C(c);
    const double *q[3] = {qu, qv, qw}, *l[3] = {lu, lv, lw};
    double *o[3] = {ou, ov, ow};
    for (int i = 0; i < 3; i++)
        if (!q[i] || !l[i] || !o[i]) return pll_set_error(PLL_EINVAL, "NULL field");
    return kl(pll_launch_jtv(&c->g, q, l, o, alpha, a0, beta, (hipStream_t)stream), "rk vjp");
}

/* mod_monitor.f90:225-258 (calcAndWriteSkf) at both walls, averaged:
 * tau = re_inv (mean_ij u(i,j,kw) + u_crf) / (dzf(kw) / 2). */
int pll_channel_wall_shear(pll_channel *c, const double *u, double *tau, void *stream)
{
    CHK_C(c);
    if (!u || !tau) return pll_set_error(PLL_EINVAL, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = kl(pll_launch_wall_sum(&c->g, u, c->part, c->scal + 6, s), "wall shear")))
        return rc;
    double sums[2];
    HIPCHK(hipMemcpyAsync(sums, c->scal + 6, sizeof sums, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *tau = pll_wall_shear_from_sums(c, sums);
    return PLL_OK;
}

/* calcAndWriteSkf, mod_monitor.f90:225-258, for each wall (k = 1 with
 * dz = dzf(1)/2, k = nz with dz = dzf(nz)/2, mod_monitor.f90:168-171):
 * tau_w = re_inv (mean_ij u(i,j,kw) + u_crf) / dz, u_tau = sqrt(tau_w),
 * cf = tau_w / (u_ref^2 / 2).  The sums come from one on-device reduction. */
int pll_channel_skin_friction(pll_channel *c, const double *u, double u_ref, double out[4],
                              void *stream)
{
    CHK_C(c);
    if (!u || !out) return pll_set_error(PLL_EINVAL, "NULL argument");
    if (!(u_ref != 0.0)) return pll_set_error(PLL_EINVAL, "u_ref must be nonzero");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = kl(pll_launch_wall_sum(&c->g, u, c->part, c->scal + 6, s), "skin friction")))
        return rc;
    double sums[2];
    HIPCHK(hipMemcpyAsync(sums, c->scal + 6, sizeof sums, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double nxy = (double)c->g.nx * c->g.ny;
    const int kw[2] = {1, c->g.nz};
    for (int w = 0; w < 2; w++) {
        const double tau = c->g.re_inv * (sums[w] / nxy + c->g.u_crf) / (0.5 * c->dzf_h[kw[w]]);
        out[2 * w] = sqrt(tau); /* NaN for a negative tau_w, as Fortran's sqrt */
        out[2 * w + 1] = tau / (0.5 * u_ref * u_ref);
    }
    return PLL_OK;
}

/* extractAndWriteProbePoint, mod_monitor.f90:146-167: u(i,j,k) + u_crf,
 * v(i,j,k), w(i,j,k), p(i,j,k) at the 1-based interior index (i,j,k). */
int pll_channel_probe(pll_channel *c, const double *u, const double *v, const double *w,
                      const double *p, int i, int j, int k, double out[4], void *stream)
{
    CHK_C(c);
    if (!u || !v || !w || !p || !out) return pll_set_error(PLL_EINVAL, "NULL argument");
    if (i < 1 || i > c->g.nx || j < 1 || j > c->g.ny || k < 1 || k > c->g.nz)
        return pll_set_error(PLL_EINVAL, "probe point (%d,%d,%d) outside 1..(%d,%d,%d)", i, j, k,
                             c->g.nx, c->g.ny, c->g.nz);
    hipStream_t s = (hipStream_t)stream;
    const size_t at = ((size_t)k * (c->g.ny + 2) + j) * (c->g.nx + 2) + i;
    const double *f[4] = {u, v, w, p};
    for (int q = 0; q < 4; q++)
        HIPCHK(hipMemcpyAsync(out + q, f[q] + at, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    out[0] += c->g.u_crf;
    return PLL_OK;
}

double pll_wall_shear_from_sums(const pll_channel *c, const double sums[2])
{
    const double nxy = (double)c->g.nx * c->g.ny;
    const double tb = c->g.re_inv * (sums[0] / nxy + c->g.u_crf) / (0.5 * c->dzf_h[1]);
    const double tt = c->g.re_inv * (sums[1] / nxy + c->g.u_crf) / (0.5 * c->dzf_h[c->g.nz]);
    return 0.5 * (tb + tt);
}
