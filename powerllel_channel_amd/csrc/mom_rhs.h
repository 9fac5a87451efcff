// Momentum right-hand side of the RK stencils, shared by the step kernels
// (scalar type double, compiled without contraction: the reference's
// rounding) and the adjoint kernels (dual numbers: its exact derivative).
#ifndef PLL_MOM_RHS_H
#define PLL_MOM_RHS_H

#include <hip/hip_runtime.h>

#include "pll_internal.h"

namespace pll {

struct G {
    int nx, ny, nz;
    double dx_inv, dy_inv, re_inv, dt, u_crf;
    const double *dzf, *dzf_inv, *dzc_inv, *dzflzi;
    long long sy, sz;  // strides of j and k
    int kc;            // planes per k chunk of the stencil tiles (0: default)
};

__host__ inline G make_g(const pll_grid_dev *g)
{
    G r;
    r.nx = g->nx;
    r.ny = g->ny;
    r.nz = g->nz;
    r.dx_inv = g->dx_inv;
    r.dy_inv = g->dy_inv;
    r.re_inv = g->re_inv;
    r.dt = g->dt;
    r.u_crf = g->u_crf;
    r.kc = g->rk_kc;
    r.dzf = g->dzf;
    r.dzf_inv = g->dzf_inv;
    r.dzc_inv = g->dzc_inv;
    r.dzflzi = g->dzflzi;
    r.sy = g->nx + 2;
    r.sz = (long long)(g->nx + 2) * (g->ny + 2);
    return r;
}

// Neighbourhood of cell (i,j,k) read by the momentum right-hand sides; the
// suffix gives the (i,j,k) offsets, p = +1, m = -1.
template <typename T>
struct NbT {
    T u000, up00, um00, u0p0, u0m0, u00p, u00m, ump0, um0p;
    T v000, vp00, vm00, v0p0, v0m0, v00p, v00m, vpm0, v0mp;
    T w000, wp00, wm00, w0p0, w0m0, w00p, w00m, wp0m, w0pm;
};
using Nb = NbT<double>;

// The z metric tables are never written by the kernels that read them; loads
// through the constant address space let the scalar unit serve them (their
// index is the same for a whole workgroup), so that they wait on the scalar
// counter and not behind the vector loads in flight.
__device__ __forceinline__ double zm(const double *p, int i)
{
    return ((const __attribute__((address_space(4))) double *)p)[i];
}

// Momentum right-hand sides (visc - conv) of the u, v and w equations at
// interior cell (i,j,k): mod_calcVel.f90:74-138, same operation order.
template <typename T>
__device__ __forceinline__ void mom_rhs(const G &g, const NbT<T> &n, int k, T &fu, T &fv, T &fw)
{
    const double dx_inv = g.dx_inv, dy_inv = g.dy_inv, re_inv = g.re_inv, u_crf = g.u_crf;
    const double dzf_km = zm(g.dzf, k - 1), dzf_k = zm(g.dzf, k), dzf_kp = zm(g.dzf, k + 1);
    const double dzfi_k = zm(g.dzf_inv, k), dzfi_kp = zm(g.dzf_inv, k + 1);
    const double dzci_k = zm(g.dzc_inv, k), dzci_km = zm(g.dzc_inv, k - 1);
    T uw, ue, us, un, ub, ut, ww, we, ws, wn, wb, wt, vw, ve, vs, vn, vb, vt;
    T dqw, dqe, dqs, dqn, dqb, dqt, conv, visc;

    ue = (n.u000 + n.up00);
    uw = (n.um00 + n.u000);
    un = (n.u000 + n.u0p0);
    us = (n.u0m0 + n.u000);
    ut = (n.u000 * dzf_kp + n.u00p * dzf_k) * dzci_k;
    ub = (n.u000 * dzf_km + n.u00m * dzf_k) * dzci_km;
    vn = (n.v000 + n.vp00);
    vs = (n.v0m0 + n.vpm0);
    wt = (n.w000 + n.wp00);
    wb = (n.w00m + n.wp0m);
    dqe = (n.up00 - n.u000) * dx_inv;
    dqw = (n.u000 - n.um00) * dx_inv;
    dqn = (n.u0p0 - n.u000) * dy_inv;
    dqs = (n.u000 - n.u0m0) * dy_inv;
    dqt = (n.u00p - n.u000) * dzci_k;
    dqb = (n.u000 - n.u00m) * dzci_km;
    conv = 0.25 * ((ue * ue - uw * uw) * dx_inv + (un * vn - us * vs) * dy_inv + (ut * wt - ub * wb) * dzfi_k);
    conv = conv + 0.5 * u_crf * (ue - uw) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzfi_k) * re_inv;
    fu = visc - conv;

    ve = (n.v000 + n.vp00);
    vw = (n.vm00 + n.v000);
    vn = (n.v000 + n.v0p0);
    vs = (n.v0m0 + n.v000);
    vt = (n.v000 * dzf_kp + n.v00p * dzf_k) * dzci_k;
    vb = (n.v000 * dzf_km + n.v00m * dzf_k) * dzci_km;
    ue = (n.u000 + n.u0p0);
    uw = (n.um00 + n.ump0);
    wt = (n.w000 + n.w0p0);
    wb = (n.w00m + n.w0pm);
    dqe = (n.vp00 - n.v000) * dx_inv;
    dqw = (n.v000 - n.vm00) * dx_inv;
    dqn = (n.v0p0 - n.v000) * dy_inv;
    dqs = (n.v000 - n.v0m0) * dy_inv;
    dqt = (n.v00p - n.v000) * dzci_k;
    dqb = (n.v000 - n.v00m) * dzci_km;
    conv = 0.25 * ((ue * ve - uw * vw) * dx_inv + (vn * vn - vs * vs) * dy_inv + (wt * vt - wb * vb) * dzfi_k);
    conv = conv + 0.5 * u_crf * (ve - vw) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzfi_k) * re_inv;
    fv = visc - conv;

    we = (n.w000 + n.wp00);
    ww = (n.wm00 + n.w000);
    wn = (n.w000 + n.w0p0);
    ws = (n.w0m0 + n.w000);
    wt = (n.w000 + n.w00p);
    wb = (n.w000 + n.w00m);
    ue = (n.u000 * dzf_kp + n.u00p * dzf_k) * dzci_k;
    uw = (n.um00 * dzf_kp + n.um0p * dzf_k) * dzci_k;
    vn = (n.v000 * dzf_kp + n.v00p * dzf_k) * dzci_k;
    vs = (n.v0m0 * dzf_kp + n.v0mp * dzf_k) * dzci_k;
    dqe = (n.wp00 - n.w000) * dx_inv;
    dqw = (n.w000 - n.wm00) * dx_inv;
    dqn = (n.w0p0 - n.w000) * dy_inv;
    dqs = (n.w000 - n.w0m0) * dy_inv;
    dqt = (n.w00p - n.w000) * dzfi_kp;
    dqb = (n.w000 - n.w00m) * dzfi_k;
    conv = 0.25 * ((ue * we - uw * ww) * dx_inv + (vn * wn - vs * ws) * dy_inv + (wt * wt - wb * wb) * dzci_k);
    conv = conv + 0.5 * u_crf * (we - ww) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzci_k) * re_inv;
    fw = visc - conv;
}

}  // namespace pll

#endif
