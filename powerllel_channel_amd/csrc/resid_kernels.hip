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

// ------------------------------------------------------------ lowest modes
// The FP32 transforms of a refinement sweep carry an absolute error of about
// 6e-8 |r| into every spectral coefficient; the lowest (kx, ky) modes have
// the smallest eigenvalues, so their solution is the least accurate in
// relative terms, and the error left after one sweep sits there.  These modes
// (kx = 0, 1 and ky = 0, 1, ny - 1) are therefore projected from the FP64
// rows in the x forward transform (low_rx), summed over y and solved here in
// FP64, removed from the FP32 spectrum, and added back in the x backward
// transform.  768x384x256, CPU emulation of the FP32 pipeline: error after
// one sweep 4.4e-11 with all modes in FP32, 7.7e-12 with these six in FP64.
namespace {

// plane k: low_y[6k + 3 kx + q] = sum_j low_rx[2 (k ny + j) + kx] exp(-2 pi i ky j / ny),
// ky = 0, 1, ny - 1 for q = 0, 1, 2; one workgroup per plane, fixed order
__global__ void __launch_bounds__(256) k_low_y(const double2 *__restrict__ rx,
                                               const double2 *__restrict__ twy,
                                               double2 *__restrict__ ly, int ny)
{
    __shared__ double red[12][256];
    const int k = blockIdx.x, t = threadIdx.x;
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) acc[q] = 0.0;
    for (int j = t; j < ny; j += 256) {
        const double2 w = twy[j];  // exp(-2 pi i j/ny)
#pragma unroll
        for (int kx = 0; kx < 2; ++kx) {
            const double2 r = rx[2 * ((long long)k * ny + j) + kx];
            acc[6 * kx + 0] += r.x;  // ky = 0
            acc[6 * kx + 1] += r.y;
            acc[6 * kx + 2] += r.x * w.x - r.y * w.y;  // ky = 1: r w
            acc[6 * kx + 3] += r.x * w.y + r.y * w.x;
            acc[6 * kx + 4] += r.x * w.x + r.y * w.y;  // ky = ny - 1: r conj(w)
            acc[6 * kx + 5] += r.y * w.x - r.x * w.y;
        }
    }
#pragma unroll
    for (int q = 0; q < 12; ++q) red[q][t] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int q = 0; q < 12; ++q) red[q][t] += red[q][t + h];
        }
        __syncthreads();
    }
    if (t < 6) ly[6 * (long long)k + t] = make_double2(red[2 * t][0], red[2 * t + 1][0]);
}

// The twelve real systems (mode q = 0..5, real and imaginary part) in one
// workgroup: the multipliers m_k = a_k / piv_{k-1} and 1 / piv_k of the four
// pivot columns (kx, |ky|) in {0,1}^2 are formed in parallel into LDS with
// the right-hand sides; then the recurrences of ssolve_trid
// (src/src_c/mod_poisson_solver.c:319-356) with the stored pivots, forward
// y_k = d_k - m_k y_{k-1}, top x = y / piv (0 for a zero pivot and for the
// pinned (0,0) mode), backward x_k = (y_k - c_k x_{k+1}) / piv_k.  Each
// system has 16 lanes of one wavefront (four systems per wavefront), lane p
// owning rows [p M, (p+1) M): both recurrences are affine in the value that
// enters a lane's rows, so a shuffle scan of the lanes' affine maps gives
// every lane its entering value, and the lane then runs the recurrence over
// its rows as written.
__global__ void __launch_bounds__(256) k_low_z(const double2 *__restrict__ ly,
                                               double2 *__restrict__ lx,
                                               const double *__restrict__ piv,
                                               const double *__restrict__ a,
                                               const double *__restrict__ c, int L, int ny, int nz,
                                               int pin_gauge)
{
    extern __shared__ double sm[];
    double *rhs = sm;                 // [nz][12]: (q, part) = 2 q + part
    double *mk = rhs + 12 * nz;       // [4][nz]  a_k / piv_{k-1}
    double *ip = mk + 4 * nz;         // [4][nz]  1 / piv_k
    double *cc = ip + 4 * nz;         // [nz]
    const long long pst = (long long)(L + 1) * (ny / 2 + 1);
    const int t = threadIdx.x;
    for (int e = t; e < 12 * nz; e += blockDim.x)
        rhs[e] = reinterpret_cast<const double *>(ly)[e];
    for (int e = t; e < 4 * nz; e += blockDim.x) {
        const int cls = e / nz, k = e - cls * nz, kx = cls & 1, iy = cls >> 1;
        const double *pc = piv + (long long)iy * (L + 1) + kx;
        const double pk = pc[(long long)k * pst];
        mk[e] = k > 0 ? a[k] / pc[(long long)(k - 1) * pst] : 0.0;
        ip[e] = pk != 0.0 ? 1.0 / pk : 0.0;
    }
    for (int k = t; k < nz; k += blockDim.x) cc[k] = c[k];
    __syncthreads();
    constexpr int P = 16;
    const int sys = t / P, p = t % P;  // sys < 12 for t < 192; lanes 192.. idle
    const int q = sys >> 1, kx = q / 3, kyq = q % 3;
    const int cls = kx + 2 * (kyq == 0 ? 0 : 1);
    const bool act = sys < 12;
    const double *m = mk + cls * nz, *r = ip + cls * nz;
    const int M = (nz + P - 1) / P;
    const int k0 = min(p * M, nz), k1 = min(k0 + M, nz);  // rows [k0, k1), maybe empty
    const int lane = t & 63;
    // forward: lane map y(k1 - 1) = al + be y(k0 - 1)
    double al = 0.0, be = 1.0;
    if (act) {
        for (int k = k0; k < k1; ++k) {
            al = rhs[12 * k + sys] - m[k] * al;
            be = -m[k] * be;
        }
    }
    for (int h = 1; h < P; h *= 2) {
        const double al2 = __shfl(al, p >= h ? lane - h : lane);
        const double be2 = __shfl(be, p >= h ? lane - h : lane);
        if (p >= h) {
            al = al + be * al2;
            be = be * be2;
        }
    }
    double y = __shfl(al, p > 0 ? lane - 1 : lane);  // y(k0 - 1); m_0 = 0 makes y(-1) irrelevant
    if (p == 0) y = 0.0;
    if (act) {
        for (int k = k0; k < k1; ++k) {
            y = rhs[12 * k + sys] - m[k] * y;
            rhs[12 * k + sys] = y;
        }
    }
    // backward: lane map x(k0) = ga + de x(k1); the top row from the top rule
    const bool pin = pin_gauge && q == 0;
    double ga = 0.0, de = 1.0;
    if (act) {
        for (int k = k1 - 1; k >= k0; --k) {
            if (k == nz - 1) {
                ga = pin ? 0.0 : rhs[12 * k + sys] * r[k];  // r = 0 for a zero pivot
                de = 0.0;
            } else {
                ga = (rhs[12 * k + sys] - cc[k] * ga) * r[k];
                de = -(cc[k] * de) * r[k];
            }
        }
    }
    for (int h = 1; h < P; h *= 2) {
        const double ga2 = __shfl(ga, p + h < P ? lane + h : lane);
        const double de2 = __shfl(de, p + h < P ? lane + h : lane);
        if (p + h < P) {
            ga = ga + de * ga2;
            de = de * de2;
        }
    }
    double x = __shfl(ga, p < P - 1 ? lane + 1 : lane);  // x(k1); unused by the top row
    if (p == P - 1) x = 0.0;
    if (act) {
        double *dst = reinterpret_cast<double *>(lx) + sys;
        for (int k = k1 - 1; k >= k0; --k) {
            x = k == nz - 1 ? (pin ? 0.0 : rhs[12 * k + sys] * r[k])
                            : (rhs[12 * k + sys] - cc[k] * x) * r[k];
            dst[12 * (long long)k] = x;
        }
    }
}

// W32 after its y forward: column 0 slot 0 real part (A_0) and slot 1 (A_1);
// column 1 slots 0, 1, ny - 1
__global__ void __launch_bounds__(256) k_low_zero(float2 *__restrict__ W, int L, int ny, int nz)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nz) return;
    float2 *P = W + (long long)k * ny * L;
    P[0].x = 0.0f;
    P[L] = make_float2(0.0f, 0.0f);
    P[1] = make_float2(0.0f, 0.0f);
    P[L + 1] = make_float2(0.0f, 0.0f);
    P[(long long)(ny - 1) * L + 1] = make_float2(0.0f, 0.0f);
}

}  // namespace

extern "C" int pll_launch_low_solve(const pll_ppe_dev *d, hipStream_t s)
{
    if (!d->low_on || d->ny < 4 || d->L < 2) return -1;
    hipLaunchKernelGGL(k_low_y, dim3(d->nz), dim3(256), 0, s, (const double2 *)d->low_rx,
                       (const double2 *)d->tw_y, (double2 *)d->low_y, d->ny);
    if (hipGetLastError() != hipSuccess) return -3;
    const size_t shm = sizeof(double) * 21 * (size_t)d->nz;
    if (shm > 163840) return -1;
    if (shm > 65536 && hipFuncSetAttribute((const void *)k_low_z,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)shm) != hipSuccess)
        return -3;
    hipLaunchKernelGGL(k_low_z, dim3(1), dim3(256), shm, s, (const double2 *)d->low_y,
                       (double2 *)d->low_x, d->piv, d->a, d->c, d->L, d->ny, d->nz, d->pin_gauge);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

extern "C" int pll_launch_low_zero(const pll_ppe_dev *d, void *W32, hipStream_t s)
{
    if (!d->low_on) return -1;
    hipLaunchKernelGGL(k_low_zero, dim3((d->nz + 255) / 256), dim3(256), 0, s, (float2 *)W32, d->L,
                       d->ny, d->nz);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

