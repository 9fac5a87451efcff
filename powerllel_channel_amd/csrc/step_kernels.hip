// RK2 projection step kernels for gfx950.
//
// Each kernel restates one Fortran routine of the reference's step, with the
// same operation order; the file is compiled with -ffp-contract=off so every
// product and sum is rounded on its own, as in the reference's source:
//   RK1 / RK2 stencils  src/mod_calcVel.f90:52-146, :240-334
//   calcRHS             src/mod_calcRHS.f90:11-37
//   correctVel          src/mod_calcVel.f90:428-453
//   forceVel + Mean_h   src/mod_calcVel.f90:455-509, src/mod_utils.f90:105-131
//   updateBoundVel/P    src/mod_updateBound.f90:30-108, :186-362 (one rank)
//   calcMaxDiv          src/mod_utils.f90:257-318
// Fields are the reference's haloed arrays f(0:nx+1,0:ny+1,0:nz+1), x fastest.
#include <hip/hip_runtime.h>

#include "pll_internal.h"

namespace {

struct G {
    int nx, ny, nz;
    double dx_inv, dy_inv, re_inv, dt, u_crf;
    const double *dzf, *dzf_inv, *dzc_inv, *dzflzi;
    long long sy, sz;  // strides of j and k
};

__host__ G make_g(const pll_grid_dev *g)
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
    r.dzf = g->dzf;
    r.dzf_inv = g->dzf_inv;
    r.dzc_inv = g->dzc_inv;
    r.dzflzi = g->dzflzi;
    r.sy = g->nx + 2;
    r.sz = (long long)(g->nx + 2) * (g->ny + 2);
    return r;
}

// Momentum right-hand sides (visc - conv) of the u, v and w equations at
// interior cell (i,j,k) for the velocity (qu,qv,qw): mod_calcVel.f90:74-138.
__device__ __forceinline__ void mom_rhs(const G &g, const double *__restrict__ qu,
                                        const double *__restrict__ qv,
                                        const double *__restrict__ qw, int i, int j, int k,
                                        double &fu, double &fv, double &fw)
{
    const long long sy = g.sy, sz = g.sz;
    const long long c = i + sy * j + sz * k;
#define U(di, dj, dk) qu[c + (di) + sy * (dj) + sz * (dk)]
#define V(di, dj, dk) qv[c + (di) + sy * (dj) + sz * (dk)]
#define W(di, dj, dk) qw[c + (di) + sy * (dj) + sz * (dk)]
    const double dx_inv = g.dx_inv, dy_inv = g.dy_inv, re_inv = g.re_inv, u_crf = g.u_crf;
    const double dzf_km = g.dzf[k - 1], dzf_k = g.dzf[k], dzf_kp = g.dzf[k + 1];
    const double dzfi_k = g.dzf_inv[k], dzfi_kp = g.dzf_inv[k + 1];
    const double dzci_k = g.dzc_inv[k], dzci_km = g.dzc_inv[k - 1];
    double uw, ue, us, un, ub, ut, ww, we, ws, wn, wb, wt, vw, ve, vs, vn, vb, vt;
    double dqw, dqe, dqs, dqn, dqb, dqt, conv, visc;

    ue = (U(0, 0, 0) + U(1, 0, 0));
    uw = (U(-1, 0, 0) + U(0, 0, 0));
    un = (U(0, 0, 0) + U(0, 1, 0));
    us = (U(0, -1, 0) + U(0, 0, 0));
    ut = (U(0, 0, 0) * dzf_kp + U(0, 0, 1) * dzf_k) * dzci_k;
    ub = (U(0, 0, 0) * dzf_km + U(0, 0, -1) * dzf_k) * dzci_km;
    vn = (V(0, 0, 0) + V(1, 0, 0));
    vs = (V(0, -1, 0) + V(1, -1, 0));
    wt = (W(0, 0, 0) + W(1, 0, 0));
    wb = (W(0, 0, -1) + W(1, 0, -1));
    dqe = (U(1, 0, 0) - U(0, 0, 0)) * dx_inv;
    dqw = (U(0, 0, 0) - U(-1, 0, 0)) * dx_inv;
    dqn = (U(0, 1, 0) - U(0, 0, 0)) * dy_inv;
    dqs = (U(0, 0, 0) - U(0, -1, 0)) * dy_inv;
    dqt = (U(0, 0, 1) - U(0, 0, 0)) * dzci_k;
    dqb = (U(0, 0, 0) - U(0, 0, -1)) * dzci_km;
    conv = 0.25 * ((ue * ue - uw * uw) * dx_inv + (un * vn - us * vs) * dy_inv + (ut * wt - ub * wb) * dzfi_k);
    conv = conv + 0.5 * u_crf * (ue - uw) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzfi_k) * re_inv;
    fu = visc - conv;

    ve = (V(0, 0, 0) + V(1, 0, 0));
    vw = (V(-1, 0, 0) + V(0, 0, 0));
    vn = (V(0, 0, 0) + V(0, 1, 0));
    vs = (V(0, -1, 0) + V(0, 0, 0));
    vt = (V(0, 0, 0) * dzf_kp + V(0, 0, 1) * dzf_k) * dzci_k;
    vb = (V(0, 0, 0) * dzf_km + V(0, 0, -1) * dzf_k) * dzci_km;
    ue = (U(0, 0, 0) + U(0, 1, 0));
    uw = (U(-1, 0, 0) + U(-1, 1, 0));
    wt = (W(0, 0, 0) + W(0, 1, 0));
    wb = (W(0, 0, -1) + W(0, 1, -1));
    dqe = (V(1, 0, 0) - V(0, 0, 0)) * dx_inv;
    dqw = (V(0, 0, 0) - V(-1, 0, 0)) * dx_inv;
    dqn = (V(0, 1, 0) - V(0, 0, 0)) * dy_inv;
    dqs = (V(0, 0, 0) - V(0, -1, 0)) * dy_inv;
    dqt = (V(0, 0, 1) - V(0, 0, 0)) * dzci_k;
    dqb = (V(0, 0, 0) - V(0, 0, -1)) * dzci_km;
    conv = 0.25 * ((ue * ve - uw * vw) * dx_inv + (vn * vn - vs * vs) * dy_inv + (wt * vt - wb * vb) * dzfi_k);
    conv = conv + 0.5 * u_crf * (ve - vw) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzfi_k) * re_inv;
    fv = visc - conv;

    we = (W(0, 0, 0) + W(1, 0, 0));
    ww = (W(-1, 0, 0) + W(0, 0, 0));
    wn = (W(0, 0, 0) + W(0, 1, 0));
    ws = (W(0, -1, 0) + W(0, 0, 0));
    wt = (W(0, 0, 0) + W(0, 0, 1));
    wb = (W(0, 0, 0) + W(0, 0, -1));
    ue = (U(0, 0, 0) * dzf_kp + U(0, 0, 1) * dzf_k) * dzci_k;
    uw = (U(-1, 0, 0) * dzf_kp + U(-1, 0, 1) * dzf_k) * dzci_k;
    vn = (V(0, 0, 0) * dzf_kp + V(0, 0, 1) * dzf_k) * dzci_k;
    vs = (V(0, -1, 0) * dzf_kp + V(0, -1, 1) * dzf_k) * dzci_k;
    dqe = (W(1, 0, 0) - W(0, 0, 0)) * dx_inv;
    dqw = (W(0, 0, 0) - W(-1, 0, 0)) * dx_inv;
    dqn = (W(0, 1, 0) - W(0, 0, 0)) * dy_inv;
    dqs = (W(0, 0, 0) - W(0, -1, 0)) * dy_inv;
    dqt = (W(0, 0, 1) - W(0, 0, 0)) * dzfi_kp;
    dqb = (W(0, 0, 0) - W(0, 0, -1)) * dzfi_k;
    conv = 0.25 * ((ue * we - uw * ww) * dx_inv + (vn * wn - vs * ws) * dy_inv + (wt * wt - wb * wb) * dzci_k);
    conv = conv + 0.5 * u_crf * (we - ww) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzci_k) * re_inv;
    fw = visc - conv;
#undef U
#undef V
#undef W
}

#define CELL_IDX                                                   \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;       \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;       \
    const int k = blockIdx.z + 1;                                  \
    if (i > g.nx || j > g.ny) return;                              \
    const long long e = i + g.sy * j + g.sz * k;

__global__ void __launch_bounds__(256) k_rk1(G g, const double *__restrict__ u,
                                              const double *__restrict__ v,
                                              const double *__restrict__ w, double *__restrict__ u1,
                                              double *__restrict__ v1, double *__restrict__ w1)
{
    CELL_IDX
    double fu, fv, fw;
    mom_rhs(g, u, v, w, i, j, k, fu, fv, fw);
    u1[e] = u[e] + g.dt * fu;
    v1[e] = v[e] + g.dt * fv;
    w1[e] = w[e] + g.dt * fw;
}

__global__ void __launch_bounds__(256) k_rk2(G g, double *__restrict__ u, double *__restrict__ v,
                                              double *__restrict__ w,
                                              const double *__restrict__ u1,
                                              const double *__restrict__ v1,
                                              const double *__restrict__ w1)
{
    CELL_IDX
    double fu, fv, fw;
    mom_rhs(g, u1, v1, w1, i, j, k, fu, fv, fw);
    u[e] = (u1[e] + g.dt * fu + u[e]) * 0.5;
    v[e] = (v1[e] + g.dt * fv + v[e]) * 0.5;
    w[e] = (w1[e] + g.dt * fw + w[e]) * 0.5;
}

__global__ void __launch_bounds__(256) k_calc_rhs(G g, const double *__restrict__ u,
                                                   const double *__restrict__ v,
                                                   const double *__restrict__ w,
                                                   double *__restrict__ rhs, double dti,
                                                   double dtidxi, double dtidyi)
{
    CELL_IDX
    rhs[e] = (u[e] - u[e - 1]) * dtidxi + (v[e] - v[e - g.sy]) * dtidyi +
             (w[e] - w[e - g.sz]) * dti * g.dzf_inv[k];
}

__global__ void __launch_bounds__(256) k_correct_vel(G g, const double *__restrict__ p,
                                                      double *__restrict__ u,
                                                      double *__restrict__ v,
                                                      double *__restrict__ w, double dtdxi,
                                                      double dtdyi)
{
    CELL_IDX
    const double pc = p[e];
    u[e] = u[e] - (p[e + 1] - pc) * dtdxi;
    v[e] = v[e] - (p[e + g.sy] - pc) * dtdyi;
    w[e] = w[e] - (p[e + g.sz] - pc) * g.dt * g.dzc_inv[k];
}

// One workgroup per interior (j,k) row: sum_i q(i,j,k)*dzflzi(k).
__global__ void __launch_bounds__(256) k_mean_part(G g, const double *__restrict__ q,
                                                    double *__restrict__ part)
{
    __shared__ double sh[256];
    const int j = blockIdx.x % g.ny + 1, k = blockIdx.x / g.ny + 1;
    const long long base = g.sy * j + g.sz * k;
    const double wk = g.dzflzi[k];
    double s = 0.0;
    for (int i = threadIdx.x + 1; i <= g.nx; i += blockDim.x) s = s + q[base + i] * wk;
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(1024) k_mean_final(int n, const double *__restrict__ part,
                                                      double *__restrict__ out, double nx,
                                                      double ny)
{
    __shared__ double sh[1024];
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) s += part[t];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] / nx / ny;
}

__global__ void __launch_bounds__(256) k_force_shift(G g, double *__restrict__ q, double vel_force,
                                                      const double *__restrict__ mean)
{
    CELL_IDX
    const double force = vel_force - mean[0];
    q[e] = q[e] + force;
}

__device__ __forceinline__ int wrap(int i, int n) { return i == 0 ? n : (i == n + 1 ? 1 : i); }

// Ghost cells after updateHalo (y periodic, full x/z range), imposePeriodicBC
// in x (full y/z range) and the wall conditions, in that order.  Every ghost
// value is a function of interior values only:
//   1 <= k <= nz:  q(wrap i, wrap j, k)        (w: 0 on the wall row k = nz)
//   k = 0, nz+1:   u,v: 2*bc - q(wrap i, wrap j, 1 | nz);  w: 0
// Work items: the two z planes (plus w's wall row k = nz), then the y faces
// and x faces of the interior k range.
__device__ __forceinline__ bool shell_pos(const G &g, long long t, int &i, int &j, int &k,
                                          int nplanes, const int *planes)
{
    const long long nxh = g.nx + 2, nyh = g.ny + 2;
    const long long nzp = nxh * nyh * nplanes;
    if (t < nzp) {
        const int pl = (int)(t / (nxh * nyh));
        const long long r = t % (nxh * nyh);
        i = (int)(r % nxh);
        j = (int)(r / nxh);
        k = planes[pl];
        return true;
    }
    t -= nzp;
    const long long nyf = 2 * nxh * g.nz;
    if (t < nyf) {
        const int side = (int)(t / (nxh * g.nz));
        const long long r = t % (nxh * g.nz);
        i = (int)(r % nxh);
        k = (int)(r / nxh) + 1;
        j = side ? g.ny + 1 : 0;
        return true;
    }
    t -= nyf;
    const long long nxf = 2LL * g.ny * g.nz;
    if (t < nxf) {
        const int side = (int)(t / ((long long)g.ny * g.nz));
        const long long r = t % ((long long)g.ny * g.nz);
        j = (int)(r % g.ny) + 1;
        k = (int)(r / g.ny) + 1;
        i = side ? g.nx + 1 : 0;
        return true;
    }
    return false;
}

__global__ void __launch_bounds__(256) k_bc_vel(G g, double *__restrict__ u, double *__restrict__ v,
                                                 double *__restrict__ w, long long total)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int planes[3] = {0, g.nz + 1, g.nz};
    int i, j, k;
    if (!shell_pos(g, t, i, j, k, 3, planes)) return;
    const long long e = i + g.sy * j + g.sz * k;
    const int ii = wrap(i, g.nx), jj = wrap(j, g.ny);
    const double bcu = 0.0 - g.u_crf;
    if (k == 0 || k == g.nz + 1) {
        const int ks = k == 0 ? 1 : g.nz;
        const long long s = ii + g.sy * jj + g.sz * ks;
        u[e] = 2.0 * bcu - u[s];
        v[e] = 2.0 * 0.0 - v[s];
        w[e] = 0.0;
    } else if (k == g.nz && t < 3LL * (g.nx + 2) * (g.ny + 2)) {
        // w's wall row, every (i,j); u and v there are ordinary cells.
        w[e] = 0.0;
        if (i != ii || j != jj) {
            const long long s = ii + g.sy * jj + g.sz * k;
            u[e] = u[s];
            v[e] = v[s];
        }
    } else {
        const long long s = ii + g.sy * jj + g.sz * k;
        u[e] = u[s];
        v[e] = v[s];
        w[e] = (k == g.nz) ? 0.0 : w[s];
    }
}

__global__ void __launch_bounds__(256) k_bc_p(G g, double *__restrict__ p, long long total)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int planes[2] = {0, g.nz + 1};
    int i, j, k;
    if (!shell_pos(g, t, i, j, k, 2, planes)) return;
    const long long e = i + g.sy * j + g.sz * k;
    const int ii = wrap(i, g.nx), jj = wrap(j, g.ny);
    const int ks = k == 0 ? 1 : (k == g.nz + 1 ? g.nz : k);
    p[e] = p[ii + g.sy * jj + g.sz * ks];
}

// One workgroup per interior (j,k) row: max |div| and the smallest linear
// index (k-1)*nx*ny + (j-1)*nx + (i-1) attaining it.
__global__ void __launch_bounds__(256) k_maxdiv_part(G g, const double *__restrict__ u,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ w,
                                                      double *__restrict__ pval,
                                                      long long *__restrict__ pidx)
{
    __shared__ double sv[256];
    __shared__ long long si[256];
    const int j = blockIdx.x % g.ny + 1, k = blockIdx.x / g.ny + 1;
    const long long base = g.sy * j + g.sz * k;
    double best = -1.0;
    long long bi = 0;
    for (int i = threadIdx.x + 1; i <= g.nx; i += blockDim.x) {
        const long long e = base + i;
        const double d = fabs((u[e] - u[e - 1]) * g.dx_inv + (v[e] - v[e - g.sy]) * g.dy_inv +
                              (w[e] - w[e - g.sz]) * g.dzf_inv[k]);
        if (d > best) {
            best = d;
            bi = (long long)(k - 1) * g.nx * g.ny + (long long)(j - 1) * g.nx + (i - 1);
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const double o = sv[threadIdx.x + h];
            const long long oi = si[threadIdx.x + h];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pval[blockIdx.x] = sv[0];
        pidx[blockIdx.x] = si[0];
    }
}

__global__ void __launch_bounds__(1024) k_maxdiv_final(int n, const double *__restrict__ pval,
                                                        const long long *__restrict__ pidx,
                                                        double *__restrict__ oval,
                                                        long long *__restrict__ oidx)
{
    __shared__ double sv[1024];
    __shared__ long long si[1024];
    double best = -1.0;
    long long bi = 0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const double o = pval[t];
        const long long oi = pidx[t];
        if (o > best || (o == best && oi < bi)) {
            best = o;
            bi = oi;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const double o = sv[threadIdx.x + h];
            const long long oi = si[threadIdx.x + h];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        oval[0] = sv[0];
        oidx[0] = si[0];
    }
}

inline dim3 cell_grid(const G &g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }
inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -3; }

}  // namespace

extern "C" int pll_launch_rk1(const pll_grid_dev *gd, const double *u, const double *v,
                              const double *w, double *u1, double *v1, double *w1, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_rk1, cell_grid(g), dim3(64, 4, 1), 0, s, g, u, v, w, u1, v1, w1);
    return ok();
}

extern "C" int pll_launch_rk2(const pll_grid_dev *gd, double *u, double *v, double *w,
                              const double *u1, const double *v1, const double *w1, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_rk2, cell_grid(g), dim3(64, 4, 1), 0, s, g, u, v, w, u1, v1, w1);
    return ok();
}

extern "C" int pll_launch_calc_rhs(const pll_grid_dev *gd, const double *u, const double *v,
                                   const double *w, double *p, hipStream_t s)
{
    const G g = make_g(gd);
    const double dti = 1.0 / g.dt;  // mod_calcRHS.f90:21-23
    const double dtidxi = dti * g.dx_inv, dtidyi = dti * g.dy_inv;
    hipLaunchKernelGGL(k_calc_rhs, cell_grid(g), dim3(64, 4, 1), 0, s, g, u, v, w, p, dti, dtidxi,
                       dtidyi);
    return ok();
}

extern "C" int pll_launch_correct_vel(const pll_grid_dev *gd, const double *p, double *u,
                                      double *v, double *w, hipStream_t s)
{
    const G g = make_g(gd);
    const double dtdxi = g.dt * g.dx_inv, dtdyi = g.dt * g.dy_inv;  // mod_calcVel.f90:436-437
    hipLaunchKernelGGL(k_correct_vel, cell_grid(g), dim3(64, 4, 1), 0, s, g, p, u, v, w, dtdxi,
                       dtdyi);
    return ok();
}

extern "C" int pll_launch_mean_partials(const pll_grid_dev *gd, const double *q, double *part,
                                        hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_mean_part, dim3(g.ny * g.nz), dim3(256), 0, s, g, q, part);
    return ok();
}

extern "C" int pll_launch_mean_final(int n, const double *part, double *out, int nx, int ny,
                                      hipStream_t s)
{
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(1024), 0, s, n, part, out, (double)nx,
                       (double)ny);
    return ok();
}

extern "C" int pll_launch_force_shift(const pll_grid_dev *gd, double *q, double vel_force,
                                      const double *mean, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_force_shift, cell_grid(g), dim3(64, 4, 1), 0, s, g, q, vel_force, mean);
    return ok();
}

extern "C" int pll_launch_bc_vel(const pll_grid_dev *gd, double *u, double *v, double *w,
                                 hipStream_t s)
{
    const G g = make_g(gd);
    const long long total = 3LL * (g.nx + 2) * (g.ny + 2) + 2LL * (g.nx + 2) * g.nz +
                            2LL * g.ny * g.nz;
    hipLaunchKernelGGL(k_bc_vel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, u, v,
                       w, total);
    return ok();
}

extern "C" int pll_launch_bc_p(const pll_grid_dev *gd, double *p, hipStream_t s)
{
    const G g = make_g(gd);
    const long long total = 2LL * (g.nx + 2) * (g.ny + 2) + 2LL * (g.nx + 2) * g.nz +
                            2LL * g.ny * g.nz;
    hipLaunchKernelGGL(k_bc_p, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, p, total);
    return ok();
}

extern "C" int pll_launch_maxdiv_partials(const pll_grid_dev *gd, const double *u,
                                          const double *v, const double *w, double *pval,
                                          long long *pidx, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_maxdiv_part, dim3(g.ny * g.nz), dim3(256), 0, s, g, u, v, w, pval, pidx);
    return ok();
}

extern "C" int pll_launch_maxdiv_final(int n, const double *pval, const long long *pidx,
                                       double *oval, long long *oidx, hipStream_t s)
{
    hipLaunchKernelGGL(k_maxdiv_final, dim3(1), dim3(1024), 0, s, n, pval, pidx, oval, oidx);
    return ok();
}
