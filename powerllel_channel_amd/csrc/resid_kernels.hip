// FP64 residual of the pressure Poisson problem for the refined solve.
//
// r = f - L_h p with L_h the operator the FFT/tridiagonal solver inverts:
//   (p_{i+1} - 2 p_i + p_{i-1}) / dx^2 + (p_{j+1} - 2 p_j + p_{j-1}) / dy^2
//   + a_k p_{k-1} + b_k p_k + c_k p_{k+1}
// with periodic i, j (the eigenvalues 2(cos(2 pi m/n) - 1)/h^2 of
// src/src_c/mod_fft_fftw.inc:180-183 are exactly those of the periodic second
// difference) and the wall-closed z coefficients of set_trid_coeff
// (src/src_c/mod_poisson_solver.c:40-103), where a_0 = c_{nz-1} = 0.
// f and p are each the interior of a haloed array or a dense array (x
// fastest); r is written in FP32 for the FP32 correction solve.
//
// Every term is formed from neighbour differences, with the z diagonal split
// as b_k = d_k - a_k - c_k (d_k the row sum, zero up to the rounding of b):
//   L_h p = ((p_{i+1}-p_i) + (p_{i-1}-p_i))/dx^2 + (same in y)
//           + a_k (p_{k-1}-p_k) + c_k (p_{k+1}-p_k) + d_k p_k.
// Differences of neighbouring values are nearly exact, so the rounding error
// of r scales with |grad p| rather than with |p| / dz^2; the plain
// a p_{k-1} + b p_k + c p_{k+1} form loses about log10(max|b| / |L_h p|)
// digits to cancellation, which on a stretched wall-refined grid caps the
// refined answer far above FP64 accuracy in the smooth (low) modes.
//
// Gauge (walls Neumann at both ends): the (0,0) mode of every solve has its
// top value pinned to 0, i.e. the plane sum of p at k = nz-1 is 0.  A
// correction only fixes equations k < nz-1 of that mode and adds a constant
// so that its own top value is 0, so the constant left by the first (FP32)
// solve, the roundoff of its top-plane mean, would stay in the answer.  The
// last sweep therefore also sums P over the top plane (per-block partials,
// then one block in fixed order) and the final write subtracts that mean.
//
// Each thread owns one (i, j) column and marches over a chunk of k, keeping
// p(k-1), p(k), p(k+1) in registers, so each p value is fetched from HBM
// once per chunk; x and y neighbours come through the caches.
#include <hip/hip_runtime.h>

#include "pll_internal.h"

namespace {

constexpr int TX = 64, TY = 4, KC = 16;

// Element (i, j, k) of a field at base + k s1 + j s2 + i: a haloed array
// (base = s1 + s2 + 1, s1 = (nx+2)(ny+2), s2 = nx+2) or a dense one (0, nx ny, nx).
struct Lay {
    long long base, s1, s2;
};

struct RArgs {
    const double *f;   // right-hand side
    const double *P;   // current solution
    Lay fl, pl;
    float *r;          // dense
    double *gpart;     // non-null: per-block sums of P over k = nz-1
    const double *a, *c, *d;
    double idx2, idy2;
    int nx, ny, nz;
};

__global__ void __launch_bounds__(TX *TY) k_resid(RArgs g)
{
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const bool valid = i < g.nx && j < g.ny;
    const int k0 = blockIdx.z * KC;
    const int k1 = min(k0 + KC, g.nz);
    double top = 0.0;
    if (valid) {
        const long long plane = (long long)g.nx * g.ny;
        const int ip = i + 1 == g.nx ? 0 : i + 1, im = i == 0 ? g.nx - 1 : i - 1;
        const int jp = j + 1 == g.ny ? 0 : j + 1, jm = j == 0 ? g.ny - 1 : j - 1;
        const long long cidx = (long long)j * g.nx + i;
        const long long s1 = g.pl.s1, s2 = g.pl.s2;
        const double *__restrict__ P = g.P + g.pl.base;  // P[k s1 + j s2 + i]
        const double *__restrict__ F = g.f + g.fl.base;
        const long long pc0 = (long long)j * s2 + i;
        double pm = k0 > 0 ? P[(k0 - 1) * s1 + pc0] : 0.0;
        double pc = P[k0 * s1 + pc0];
        for (int k = k0; k < k1; ++k) {
            const long long kb = (long long)k * s1;
            const double pp = k + 1 < g.nz ? P[kb + s1 + pc0] : 0.0;
            const double lx = ((P[kb + (long long)j * s2 + ip] - pc) + (P[kb + (long long)j * s2 + im] - pc)) * g.idx2;
            const double ly = ((P[kb + (long long)jp * s2 + i] - pc) + (P[kb + (long long)jm * s2 + i] - pc)) * g.idy2;
            const double lz = g.a[k] * (pm - pc) + g.c[k] * (pp - pc) + g.d[k] * pc;
            const double f = F[k * g.fl.s1 + (long long)j * g.fl.s2 + i];
            g.r[k * plane + cidx] = (float)(f - (lx + ly + lz));
            if (k == g.nz - 1) top = pc;
            pm = pc;
            pc = pp;
        }
    }
    if (g.gpart && k1 == g.nz) {  // uniform over the block
        __shared__ double red[TX * TY];
        const int t = threadIdx.y * TX + threadIdx.x;
        red[t] = top;
        __syncthreads();
        for (int h = TX * TY / 2; h > 0; h >>= 1) {
            if (t < h) red[t] += red[t + h];
            __syncthreads();
        }
        if (t == 0) g.gpart[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
    }
}

// The top-plane partial sums of k_resid alone (same blocks, same tree): for
// the forms that fuse the residual into the x transform.
__global__ void __launch_bounds__(TX *TY) k_top_parts(const double *__restrict__ P, Lay pl,
                                                      int nx, int ny, int nz,
                                                      double *__restrict__ gpart)
{
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const double top =
        (i < nx && j < ny) ? P[pl.base + (long long)(nz - 1) * pl.s1 + (long long)j * pl.s2 + i] : 0.0;
    __shared__ double red[TX * TY];
    const int t = threadIdx.y * TX + threadIdx.x;
    red[t] = top;
    __syncthreads();
    for (int h = TX * TY / 2; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) gpart[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// eps = (sum of the n partials, in a fixed order) * inv
__global__ void __launch_bounds__(256) k_gauge_final(const double *part, int n, double inv,
                                                     double *eps)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int q = t; q < n; q += 256) s += part[q];
    red[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) eps[0] = red[0] * inv;
}

}  // namespace

extern "C" int pll_gauge_parts(const pll_ppe_dev *d)
{
    return ((d->nx + TX - 1) / TX) * ((d->ny + TY - 1) / TY);
}

static Lay halo_lay(const pll_ppe_dev *d)
{
    const long long s1 = (long long)(d->nx + 2) * (d->ny + 2), s2 = d->nx + 2;
    return Lay{s1 + s2 + 1, s1, s2};
}

static Lay dense_lay(const pll_ppe_dev *d)
{
    return Lay{0, (long long)d->nx * d->ny, d->nx};
}

static int launch_resid(const pll_ppe_dev *d, const double *f, Lay fl, const double *P, Lay pl,
                        float *r, double *gauge, hipStream_t s)
{
    RArgs g;
    g.f = f;
    g.P = P;
    g.fl = fl;
    g.pl = pl;
    g.r = r;
    g.gpart = gauge;  // gauge[0 .. parts): partials, gauge[parts]: the mean
    g.a = d->a;
    g.c = d->c;
    g.d = d->dz;
    g.idx2 = d->idx2;
    g.idy2 = d->idy2;
    g.nx = d->nx;
    g.ny = d->ny;
    g.nz = d->nz;
    const dim3 grid((d->nx + TX - 1) / TX, (d->ny + TY - 1) / TY, (d->nz + KC - 1) / KC);
    hipLaunchKernelGGL(k_resid, grid, dim3(TX, TY), 0, s, g);
    if (hipGetLastError() != hipSuccess) return -3;
    if (gauge) {
        const int n = pll_gauge_parts(d);
        hipLaunchKernelGGL(k_gauge_final, dim3(1), dim3(256), 0, s, gauge, n,
                           1.0 / ((double)d->nx * d->ny), gauge + n);
        if (hipGetLastError() != hipSuccess) return -3;
    }
    return 0;
}

extern "C" int pll_launch_resid(const pll_ppe_dev *d, const double *f_halo, const double *P,
                                float *r, double *gauge, hipStream_t s)
{
    return launch_resid(d, f_halo, halo_lay(d), P, dense_lay(d), r, gauge, s);
}

extern "C" int pll_launch_resid_hp(const pll_ppe_dev *d, const double *F, const double *p_halo,
                                   float *r, double *gauge, hipStream_t s)
{
    return launch_resid(d, F, dense_lay(d), p_halo, halo_lay(d), r, gauge, s);
}

/* gauge[parts] := mean of p over the top plane, p haloed (halo != 0) or dense */
extern "C" int pll_launch_gauge(const pll_ppe_dev *d, const double *P, int halo, double *gauge,
                                hipStream_t s)
{
    const dim3 grid((d->nx + TX - 1) / TX, (d->ny + TY - 1) / TY);
    hipLaunchKernelGGL(k_top_parts, grid, dim3(TX, TY), 0, s, P, halo ? halo_lay(d) : dense_lay(d),
                       d->nx, d->ny, d->nz, gauge);
    if (hipGetLastError() != hipSuccess) return -3;
    const int n = pll_gauge_parts(d);
    hipLaunchKernelGGL(k_gauge_final, dim3(1), dim3(256), 0, s, gauge, n,
                       1.0 / ((double)d->nx * d->ny), gauge + n);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
