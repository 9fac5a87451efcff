/*
 * Host side of the RK2 projection step on the GPU (C, drives the HIP kernels).
 *
 * One handle holds the grid (mod_mesh.f90:26-84 derived arrays on the
 * device), the physics constants and the small reduction buffers.  The
 * entry points mirror the reference's per-step routines (src/main.f90:131-212).
 */
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
    double *scal;      /* device: 4 scalars (mean, max value)  */
    long long *sidx;   /* device: 1 index */
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
    for (int d = 0; d < 3; d++) {
        c->is_forced[d] = is_forced ? is_forced[d] != 0 : (d == 0);
        c->vel_force[d] = vel_force ? vel_force[d] : (d == 0 ? 1.0 : 0.0);
    }
    hipError_t e;
    const size_t nrow = (size_t)ny * nz;
    if ((e = hipMalloc((void **)&c->zarr, sizeof(double) * 4 * (size_t)n2)) != hipSuccess ||
        (e = hipMalloc((void **)&c->part, sizeof(double) * nrow)) != hipSuccess ||
        (e = hipMalloc((void **)&c->pidx, sizeof(long long) * nrow)) != hipSuccess ||
        (e = hipMalloc((void **)&c->scal, sizeof(double) * 4)) != hipSuccess ||
        (e = hipMalloc((void **)&c->sidx, sizeof(long long))) != hipSuccess ||
        (e = hipMemcpy(c->zarr, z, sizeof(double) * 4 * (size_t)n2, hipMemcpyHostToDevice)) !=
            hipSuccess) {
        free(z);
        pll_channel_destroy(c);
        return pll_set_error(PLL_ENOMEM, "device allocation/copy failed: %s", hipGetErrorString(e));
    }
    free(z);
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
    return kl(pll_launch_rk1(&c->g, u, v, w, u1, v1, w1, (hipStream_t)stream), "rk1");
}

int pll_channel_rk2(pll_channel *c, double *u, double *v, double *w, const double *u1,
                    const double *v1, const double *w1, void *stream)
{
    CHK_C(c);
    return kl(pll_launch_rk2(&c->g, u, v, w, u1, v1, w1, (hipStream_t)stream), "rk2");
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
    return kl(pll_launch_correct_vel(&c->g, p, u, v, w, (hipStream_t)stream), "correct_vel");
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

int pll_channel_step(pll_channel *c, pll_ppe *ppe, double *u, double *v, double *w, double *u1,
                     double *v1, double *w1, double *p, void *stream)
{
    CHK_C(c);
    if (!ppe) return pll_set_error(PLL_ESTATE, "PPE handle is NULL");
    int rc;
    if ((rc = pll_channel_rk1(c, u, v, w, u1, v1, w1, stream))) return rc;
    if ((rc = pll_channel_bc_vel(c, u1, v1, w1, stream))) return rc;
    if ((rc = pll_channel_rk2(c, u, v, w, u1, v1, w1, stream))) return rc;
    if ((rc = pll_channel_bc_vel(c, u, v, w, stream))) return rc;
    if ((rc = pll_channel_calc_rhs(c, u, v, w, p, stream))) return rc;
    if ((rc = pll_ppe_solve(ppe, p, stream))) return rc;
    if ((rc = pll_channel_bc_p(c, p, stream))) return rc;
    if ((rc = pll_channel_correct_vel(c, p, u, v, w, stream))) return rc;
    if ((rc = pll_channel_force_vel(c, u, v, w, stream))) return rc;
    return pll_channel_bc_vel(c, u, v, w, stream);
}
