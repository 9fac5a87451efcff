/*
 * Adjoint of the RK2 projection step over a window of steps (no counterpart
 * in the reference, which has no adjoint or body force).
 *
 * Objective: J = (1/nsteps) sum_{n=1..nsteps} tau_w(u^n), tau_w the wall
 * shear averaged over both walls (src/mod_monitor.f90:225-258).  Control: a
 * body force b = (bu, bv, bw) added to the momentum right-hand side of both
 * RK stages of every step (pll_channel_step_body).
 *
 * Forward: the steps run with checkpoints of (u,v,w) every `ckpt_every`
 * steps.  Backward: per segment, the states of the segment are recomputed
 * from its checkpoint, then the steps are reversed; each reversed step
 * recomputes its RK1 stage (stage recomputation).  One reversed step is the
 * transpose of the forward step's linearisation, in reverse order:
 *   boundary fill, forcing, correctVel, updateBoundP, PPE (transposed solve),
 *   calcRHS, boundary fill, RK2 (Jacobian at u1), boundary fill, RK1
 *   (Jacobian at u),
 * with grad b += dt/2 l* (RK2) + dt l1 (RK1).
 */
#include <stdlib.h>
#include <string.h>

#include "../../include/powerllel.h"
#include "pll_internal.h"

struct pll_adjoint {
    pll_channel *ch;
    pll_ppe *ppe;
    int nsteps, every, nck;
    size_t fbytes;           /* bytes of one haloed field */
    double *ck;              /* nck checkpoints x 3 fields */
    double *seg;             /* every states x 3 fields */
    double *u1, *p;          /* 3 fields, 1 field */
    double *lam, *lam1, *lamp, *grad, *body; /* 3, 3, 1, 3, 3 fields */
    double *wsum;            /* device: 2 nsteps wall sums */
    double *part;            /* device: partials for the wall sums */
    double *hsum;            /* host: 2 nsteps */
    int have_forward;
    size_t dev_bytes;
};

static int hipchk(hipError_t e, const char *what)
{
    if (e == hipSuccess) return PLL_OK;
    return pll_set_error(PLL_EHIP, "%s: %s", what, hipGetErrorString(e));
}

static int launch(int rc, const char *what)
{
    if (rc == 0) return PLL_OK;
    return pll_set_error(PLL_EHIP, "%s: kernel launch failed: %s", what,
                         hipGetErrorString(hipGetLastError()));
}

int pll_adjoint_destroy(pll_adjoint *a)
{
    if (!a) return PLL_OK;
    void *bufs[] = {a->ck, a->seg, a->u1, a->p, a->lam, a->lam1, a->lamp, a->grad, a->body,
                    a->wsum, a->part};
    for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; i++)
        if (bufs[i]) hipFree(bufs[i]);
    free(a->hsum);
    free(a);
    return PLL_OK;
}

static int dalloc(pll_adjoint *a, double **p, size_t bytes)
{
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = NULL;
        return pll_set_error(PLL_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    a->dev_bytes += bytes;
    return PLL_OK;
}

int pll_adjoint_create(pll_adjoint **out, pll_channel *ch, pll_ppe *ppe, int nsteps,
                       int ckpt_every)
{
    if (!out) return pll_set_error(PLL_EINVAL, "out is NULL");
    *out = NULL;
    if (!ch || !ppe) return pll_set_error(PLL_ESTATE, "channel or PPE handle is NULL");
    if (nsteps < 1 || ckpt_every < 1)
        return pll_set_error(PLL_EINVAL, "nsteps (%d) and ckpt_every (%d) must be >= 1", nsteps,
                             ckpt_every);
    if (pll_ppe_refinement(ppe) >= 0)
        return pll_set_error(PLL_EUNSUP, "the adjoint uses the FP64 solve (refinement is on)");
    const pll_grid_dev *g = pll_channel_grid(ch);
    pll_adjoint *a = (pll_adjoint *)calloc(1, sizeof *a);
    if (!a) return pll_set_error(PLL_ENOMEM, "calloc failed");
    a->ch = ch;
    a->ppe = ppe;
    a->nsteps = nsteps;
    a->every = ckpt_every < nsteps ? ckpt_every : nsteps;
    a->nck = (nsteps + a->every - 1) / a->every;
    const size_t nf = (size_t)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
    a->fbytes = nf * sizeof(double);
    int rc;
    if ((rc = dalloc(a, &a->ck, 3 * a->fbytes * (size_t)a->nck)) ||
        (rc = dalloc(a, &a->seg, 3 * a->fbytes * (size_t)a->every)) ||
        (rc = dalloc(a, &a->u1, 3 * a->fbytes)) || (rc = dalloc(a, &a->p, a->fbytes)) ||
        (rc = dalloc(a, &a->lam, 3 * a->fbytes)) || (rc = dalloc(a, &a->lam1, 3 * a->fbytes)) ||
        (rc = dalloc(a, &a->lamp, a->fbytes)) || (rc = dalloc(a, &a->grad, 3 * a->fbytes)) ||
        (rc = dalloc(a, &a->body, 3 * a->fbytes)) ||
        (rc = dalloc(a, &a->wsum, sizeof(double) * 2 * (size_t)nsteps)) ||
        (rc = dalloc(a, &a->part, sizeof(double) * 2 * (size_t)g->ny))) {
        pll_adjoint_destroy(a);
        return rc;
    }
    a->hsum = (double *)malloc(sizeof(double) * 2 * (size_t)nsteps);
    if (!a->hsum) {
        pll_adjoint_destroy(a);
        return pll_set_error(PLL_ENOMEM, "malloc failed");
    }
    *out = a;
    return PLL_OK;
}

size_t pll_adjoint_device_bytes(const pll_adjoint *a) { return a ? a->dev_bytes : 0; }

/* field f (0..2) of the 3-field block at base */
static double *fld(const pll_adjoint *a, double *base, int f)
{
    return base + (size_t)f * (a->fbytes / sizeof(double));
}

static int copy3(const pll_adjoint *a, double *dst, const double *const src[3], hipStream_t s)
{
    for (int f = 0; f < 3; f++) {
        int rc = hipchk(hipMemcpyAsync(fld(a, dst, f), src[f], a->fbytes,
                                       hipMemcpyDeviceToDevice, s),
                        "copy");
        if (rc) return rc;
    }
    return PLL_OK;
}

static int step_fwd(pll_adjoint *a, double *U, hipStream_t s)
{
    const double *body[3] = {fld(a, a->body, 0), fld(a, a->body, 1), fld(a, a->body, 2)};
    return pll_channel_step_body_internal(a->ch, a->ppe, fld(a, U, 0), fld(a, U, 1),
                                          fld(a, U, 2), fld(a, a->u1, 0), fld(a, a->u1, 1),
                                          fld(a, a->u1, 2), a->p, body, s);
}

int pll_adjoint_forward(pll_adjoint *a, const double *u0, const double *v0, const double *w0,
                        const double *bu, const double *bv, const double *bw, double *J,
                        void *stream)
{
    if (!a) return pll_set_error(PLL_ESTATE, "adjoint handle is NULL");
    if (!u0 || !v0 || !w0 || !bu || !bv || !bw || !J)
        return pll_set_error(PLL_EINVAL, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const pll_grid_dev *g = pll_channel_grid(a->ch);
    const double *U0[3] = {u0, v0, w0}, *B[3] = {bu, bv, bw};
    int rc;
    if ((rc = copy3(a, a->body, B, s))) return rc;
    if ((rc = copy3(a, a->ck, U0, s))) return rc;
    /* the running state lives in seg[0] */
    double *S = a->seg;
    const double *C0[3] = {fld(a, a->ck, 0), fld(a, a->ck, 1), fld(a, a->ck, 2)};
    if ((rc = copy3(a, S, C0, s))) return rc;
    for (int n = 0; n < a->nsteps; n++) {
        if ((rc = step_fwd(a, S, s))) return rc;
        if ((rc = launch(pll_launch_wall_sum(g, fld(a, S, 0), a->part, a->wsum + 2 * n, s),
                         "wall shear")))
            return rc;
        if ((n + 1) % a->every == 0 && (n + 1) < a->nsteps) {
            const double *Sf[3] = {fld(a, S, 0), fld(a, S, 1), fld(a, S, 2)};
            if ((rc = copy3(a, a->ck + (size_t)((n + 1) / a->every) * 3 * (a->fbytes / 8), Sf, s)))
                return rc;
        }
    }
    if ((rc = hipchk(hipMemcpyAsync(a->hsum, a->wsum, sizeof(double) * 2 * (size_t)a->nsteps,
                                    hipMemcpyDeviceToHost, s),
                     "copy of wall sums")))
        return rc;
    if ((rc = hipchk(hipStreamSynchronize(s), "synchronise"))) return rc;
    double acc = 0.0;
    for (int n = 0; n < a->nsteps; n++) acc += pll_wall_shear_from_sums(a->ch, a->hsum + 2 * n);
    *J = acc / a->nsteps;
    a->have_forward = 1;
    return PLL_OK;
}

/* dJ/du on the wall planes for one state: (1/nsteps) * 1/2 * re_inv /
 * (nx ny dzf(kw)/2) */
static int seed(pll_adjoint *a, hipStream_t s)
{
    const pll_grid_dev *g = pll_channel_grid(a->ch);
    const double *dzf = pll_channel_dzf_host(a->ch);
    const double nxy = (double)g->nx * g->ny;
    const double cb = g->re_inv / (a->nsteps * nxy * dzf[1]);
    const double ct = g->re_inv / (a->nsteps * nxy * dzf[g->nz]);
    return launch(pll_launch_wall_seed(g, fld(a, a->lam, 0), cb, ct, s), "wall seed");
}

/* lam (adjoint of the state after step n) -> adjoint of the state before it;
 * U is that state, its RK1 stage is recomputed into u1. */
static int step_adj(pll_adjoint *a, double *U, hipStream_t s)
{
    pll_channel *ch = a->ch;
    const pll_grid_dev *g = pll_channel_grid(ch);
    double *L[3] = {fld(a, a->lam, 0), fld(a, a->lam, 1), fld(a, a->lam, 2)};
    double *L1[3] = {fld(a, a->lam1, 0), fld(a, a->lam1, 1), fld(a, a->lam1, 2)};
    double *Gr[3] = {fld(a, a->grad, 0), fld(a, a->grad, 1), fld(a, a->grad, 2)};
    const double *Uc[3] = {fld(a, U, 0), fld(a, U, 1), fld(a, U, 2)};
    double *U1[3] = {fld(a, a->u1, 0), fld(a, a->u1, 1), fld(a, a->u1, 2)};
    const double *B[3] = {fld(a, a->body, 0), fld(a, a->body, 1), fld(a, a->body, 2)};
    int rc;
    /* stage recomputation: u1 = RK1(u) with its boundary fill */
    if ((rc = launch(pll_launch_rk1(g, Uc[0], Uc[1], Uc[2], U1[0], U1[1], U1[2], B, s),
                     "rk1")))
        return rc;
    if ((rc = pll_channel_bc_vel(ch, U1[0], U1[1], U1[2], s))) return rc;
    /* final boundary fill, forcing, correctVel (velocity part is the identity) */
    if ((rc = pll_channel_bc_vel_adj(ch, L[0], L[1], L[2], s))) return rc;
    if ((rc = pll_channel_force_vel_adj(ch, L[0], L[1], L[2], s))) return rc;
    if ((rc = pll_channel_correct_vel_adj(ch, L[0], L[1], L[2], a->lamp, s))) return rc;
    if ((rc = pll_channel_bc_p_adj(ch, a->lamp, s))) return rc;
    if ((rc = pll_ppe_solve_adjoint(a->ppe, a->lamp, s))) return rc;
    if ((rc = pll_channel_calc_rhs_adj(ch, a->lamp, L[0], L[1], L[2], s))) return rc;
    if ((rc = pll_channel_bc_vel_adj(ch, L[0], L[1], L[2], s))) return rc;
    /* RK2: u* = (u1 + dt (f(u1) + b) + u) / 2 */
    if ((rc = launch(pll_launch_axpy3(g, 0.5 * g->dt, (const double *const *)L, Gr, s), "axpy")))
        return rc;
    const double *U1c[3] = {U1[0], U1[1], U1[2]};
    const double *Lc[3] = {L[0], L[1], L[2]};
    if ((rc = launch(pll_launch_jtv(g, U1c, Lc, L1, 0.5 * g->dt, 0.5, 0.0, s), "rk2 adjoint")))
        return rc;
    if ((rc = pll_channel_bc_vel_adj(ch, L1[0], L1[1], L1[2], s))) return rc;
    /* RK1: u1 = u + dt (f(u) + b) */
    if ((rc = launch(pll_launch_axpy3(g, g->dt, (const double *const *)L1, Gr, s), "axpy")))
        return rc;
    const double *L1c[3] = {L1[0], L1[1], L1[2]};
    return launch(pll_launch_jtv(g, Uc, L1c, L, g->dt, 1.0, 0.5, s), "rk1 adjoint");
}

int pll_adjoint_backward(pll_adjoint *a, double *gu, double *gv, double *gw, void *stream)
{
    if (!a) return pll_set_error(PLL_ESTATE, "adjoint handle is NULL");
    if (!a->have_forward) return pll_set_error(PLL_ESTATE, "pll_adjoint_forward not run");
    if (!gu || !gv || !gw) return pll_set_error(PLL_EINVAL, "NULL gradient field");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = hipchk(hipMemsetAsync(a->grad, 0, 3 * a->fbytes, s), "clear")) ||
        (rc = hipchk(hipMemsetAsync(a->lam, 0, 3 * a->fbytes, s), "clear")) ||
        (rc = hipchk(hipMemsetAsync(a->lam1, 0, 3 * a->fbytes, s), "clear")))
        return rc;
    if ((rc = seed(a, s))) return rc;  /* state after the last step */
    const size_t f3 = 3 * (a->fbytes / sizeof(double));
    for (int m = a->nck - 1; m >= 0; m--) {
        const int n0 = m * a->every;
        const int n1 = n0 + a->every < a->nsteps ? n0 + a->every : a->nsteps;
        /* states n0 .. n1-1 of the segment into seg[0 .. n1-n0) */
        const double *C[3] = {a->ck + (size_t)m * f3, a->ck + (size_t)m * f3 + f3 / 3,
                              a->ck + (size_t)m * f3 + 2 * (f3 / 3)};
        if ((rc = copy3(a, a->seg, C, s))) return rc;
        for (int n = n0; n < n1 - 1; n++) {
            double *src = a->seg + (size_t)(n - n0) * f3, *dst = src + f3;
            const double *Sf[3] = {src, src + f3 / 3, src + 2 * (f3 / 3)};
            if ((rc = copy3(a, dst, Sf, s))) return rc;
            if ((rc = step_fwd(a, dst, s))) return rc;
        }
        for (int n = n1 - 1; n >= n0; n--) {
            if ((rc = step_adj(a, a->seg + (size_t)(n - n0) * f3, s))) return rc;
            if (n >= 1 && (rc = seed(a, s))) return rc;  /* J reads the state n too */
        }
    }
    double *G[3] = {gu, gv, gw};
    for (int f = 0; f < 3; f++)
        if ((rc = hipchk(hipMemcpyAsync(G[f], fld(a, a->grad, f), a->fbytes,
                                        hipMemcpyDeviceToDevice, s),
                         "copy of the gradient")))
            return rc;
    return PLL_OK;
}
