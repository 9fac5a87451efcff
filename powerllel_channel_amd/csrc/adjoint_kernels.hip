// Transposes of the step's linear(ised) pieces, for the adjoint of the RK2
// projection step (no counterpart in the reference, which has no adjoint).
//
// Every kernel here writes each output element from one thread (gathers, no
// atomics), so results are deterministic.  The momentum right-hand side is
// differentiated by evaluating the step kernels' own mom_rhs on dual numbers,
// so the derivative follows the forward arithmetic exactly.
//
// Layout: all fields are the haloed arrays f(0:nx+1, 0:ny+1, 0:nz+1).
#include <hip/hip_runtime.h>

#include "mom_rhs.h"
#include "pll_internal.h"

namespace {

using pll::G;
using pll::make_g;
using pll::mom_rhs;
using pll::NbT;

// Dual number with three tangents (one per velocity component).
struct D3 {
    double v, d[3];
};
__device__ __forceinline__ D3 operator+(const D3 &a, const D3 &b)
{
    return D3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ __forceinline__ D3 operator-(const D3 &a, const D3 &b)
{
    return D3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ __forceinline__ D3 operator*(const D3 &a, const D3 &b)
{
    return D3{a.v * b.v,
              {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1],
               a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ __forceinline__ D3 operator*(const D3 &a, double s)
{
    return D3{a.v * s, {a.d[0] * s, a.d[1] * s, a.d[2] * s}};
}
__device__ __forceinline__ D3 operator*(double s, const D3 &a) { return a * s; }

// Dual number with one tangent (directional derivative).
struct D1 {
    double v, d;
};
__device__ __forceinline__ D1 operator+(const D1 &a, const D1 &b) { return D1{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D1 operator-(const D1 &a, const D1 &b) { return D1{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D1 operator*(const D1 &a, const D1 &b)
{
    return D1{a.v * b.v, a.d * b.v + a.v * b.d};
}
__device__ __forceinline__ D1 operator*(const D1 &a, double s) { return D1{a.v * s, a.d * s}; }
__device__ __forceinline__ D1 operator*(double s, const D1 &a) { return a * s; }

__device__ __forceinline__ bool interior(const G &g, int i, int j, int k)
{
    return i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz;
}

__device__ __forceinline__ bool cell_of(const G &g, long long t, int &i, int &j, int &k)
{
    const long long nxh = g.nx + 2, nyh = g.ny + 2;
    if (t >= nxh * nyh * (g.nz + 2)) return false;
    i = (int)(t % nxh);
    t /= nxh;
    j = (int)(t % nyh);
    k = (int)(t / nyh);
    return true;
}

// ------------------------------------------------------------ momentum RHS
// Tangent seed of slot offset (a,b,c) of field F when the slot's input is the
// cell c0 and the output is o = c0 + (DX,DY,DZ): the slot reads o + (a,b,c),
// which is c0 iff (a,b,c) = -(DX,DY,DZ).
template <int DX, int DY, int DZ>
__device__ __forceinline__ D3 seed(double val, int F, int a, int b, int c)
{
    const bool hit = (a == -DX && b == -DY && c == -DZ);
    return D3{val, {(hit && F == 0) ? 1.0 : 0.0, (hit && F == 1) ? 1.0 : 0.0,
                    (hit && F == 2) ? 1.0 : 0.0}};
}

// acc[F] += sum_e d f_e(o) / d q_F(c0) * l_e(o), o = c0 + (DX,DY,DZ) interior.
template <int DX, int DY, int DZ>
__device__ __forceinline__ void jt_term(const G &g, const double *__restrict__ qu,
                                        const double *__restrict__ qv,
                                        const double *__restrict__ qw,
                                        const double *__restrict__ lu,
                                        const double *__restrict__ lv,
                                        const double *__restrict__ lw, int i, int j, int k,
                                        double acc[3])
{
    const int io = i + DX, jo = j + DY, ko = k + DZ;
    if (!interior(g, io, jo, ko)) return;
    const long long sy = g.sy, sz = g.sz;
    const long long e = io + sy * jo + sz * ko;
#define S(name, arr, F, a, b, c) n.name = seed<DX, DY, DZ>(arr[e + (a) + sy * (b) + sz * (c)], F, a, b, c)
    NbT<D3> n;
    S(u000, qu, 0, 0, 0, 0); S(up00, qu, 0, 1, 0, 0); S(um00, qu, 0, -1, 0, 0);
    S(u0p0, qu, 0, 0, 1, 0); S(u0m0, qu, 0, 0, -1, 0); S(u00p, qu, 0, 0, 0, 1);
    S(u00m, qu, 0, 0, 0, -1); S(ump0, qu, 0, -1, 1, 0); S(um0p, qu, 0, -1, 0, 1);
    S(v000, qv, 1, 0, 0, 0); S(vp00, qv, 1, 1, 0, 0); S(vm00, qv, 1, -1, 0, 0);
    S(v0p0, qv, 1, 0, 1, 0); S(v0m0, qv, 1, 0, -1, 0); S(v00p, qv, 1, 0, 0, 1);
    S(v00m, qv, 1, 0, 0, -1); S(vpm0, qv, 1, 1, -1, 0); S(v0mp, qv, 1, 0, -1, 1);
    S(w000, qw, 2, 0, 0, 0); S(wp00, qw, 2, 1, 0, 0); S(wm00, qw, 2, -1, 0, 0);
    S(w0p0, qw, 2, 0, 1, 0); S(w0m0, qw, 2, 0, -1, 0); S(w00p, qw, 2, 0, 0, 1);
    S(w00m, qw, 2, 0, 0, -1); S(wp0m, qw, 2, 1, 0, -1); S(w0pm, qw, 2, 0, 1, -1);
#undef S
    D3 fu, fv, fw;
    mom_rhs(g, n, ko, fu, fv, fw);
    const double a = lu[e], b = lv[e], c = lw[e];
#pragma unroll
    for (int F = 0; F < 3; ++F) acc[F] += fu.d[F] * a + fv.d[F] * b + fw.d[F] * c;
}

// o_F(c) = beta o_F(c) + a0 l_F(c) [c interior] + alpha (J_f(q)^T l)_F(c)
// over every cell c of the haloed arrays; l is read on the interior only.
// The 13 outputs that read c: the slot offsets of NbT, negated.
__global__ void __launch_bounds__(256) k_jtv(G g, const double *__restrict__ qu,
                                             const double *__restrict__ qv,
                                             const double *__restrict__ qw,
                                             const double *__restrict__ lu,
                                             const double *__restrict__ lv,
                                             const double *__restrict__ lw,
                                             double *__restrict__ ou, double *__restrict__ ov,
                                             double *__restrict__ ow, double alpha, double a0,
                                             double beta)
{
    int i, j, k;
    if (!cell_of(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    double acc[3] = {0.0, 0.0, 0.0};
    jt_term<0, 0, 0>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);
    jt_term<-1, 0, 0>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);
    jt_term<1, 0, 0>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);
    jt_term<0, -1, 0>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);
    jt_term<0, 1, 0>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);
    jt_term<0, 0, -1>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);
    jt_term<0, 0, 1>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);
    jt_term<1, -1, 0>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);   // u: ump0
    jt_term<1, 0, -1>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);   // u: um0p
    jt_term<-1, 1, 0>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);   // v: vpm0
    jt_term<0, 1, -1>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);   // v: v0mp
    jt_term<-1, 0, 1>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);   // w: wp0m
    jt_term<0, -1, 1>(g, qu, qv, qw, lu, lv, lw, i, j, k, acc);   // w: w0pm
    const long long e = i + g.sy * j + g.sz * k;
    const bool in = interior(g, i, j, k);
    double r[3] = {alpha * acc[0], alpha * acc[1], alpha * acc[2]};
    if (in && a0 != 0.0) {
        r[0] += a0 * lu[e];
        r[1] += a0 * lv[e];
        r[2] += a0 * lw[e];
    }
    if (beta != 0.0) {
        r[0] += beta * ou[e];
        r[1] += beta * ov[e];
        r[2] += beta * ow[e];
    }
    ou[e] = r[0];
    ov[e] = r[1];
    ow[e] = r[2];
}

// o = J_f(q) dq on the interior (directional derivative of the momentum
// right-hand side), for checks of the transpose.
__global__ void __launch_bounds__(256) k_jvp(G g, const double *__restrict__ qu,
                                             const double *__restrict__ qv,
                                             const double *__restrict__ qw,
                                             const double *__restrict__ du,
                                             const double *__restrict__ dv,
                                             const double *__restrict__ dw,
                                             double *__restrict__ ou, double *__restrict__ ov,
                                             double *__restrict__ ow)
{
    int i, j, k;
    if (!cell_of(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    if (!interior(g, i, j, k)) return;
    const long long sy = g.sy, sz = g.sz;
    const long long e = i + sy * j + sz * k;
#define S(name, arr, darr, a, b, c)                               \
    do {                                                          \
        const long long x_ = e + (a) + sy * (b) + sz * (c);       \
        n.name = D1{arr[x_], darr[x_]};                           \
    } while (0)
    NbT<D1> n;
    S(u000, qu, du, 0, 0, 0); S(up00, qu, du, 1, 0, 0); S(um00, qu, du, -1, 0, 0);
    S(u0p0, qu, du, 0, 1, 0); S(u0m0, qu, du, 0, -1, 0); S(u00p, qu, du, 0, 0, 1);
    S(u00m, qu, du, 0, 0, -1); S(ump0, qu, du, -1, 1, 0); S(um0p, qu, du, -1, 0, 1);
    S(v000, qv, dv, 0, 0, 0); S(vp00, qv, dv, 1, 0, 0); S(vm00, qv, dv, -1, 0, 0);
    S(v0p0, qv, dv, 0, 1, 0); S(v0m0, qv, dv, 0, -1, 0); S(v00p, qv, dv, 0, 0, 1);
    S(v00m, qv, dv, 0, 0, -1); S(vpm0, qv, dv, 1, -1, 0); S(v0mp, qv, dv, 0, -1, 1);
    S(w000, qw, dw, 0, 0, 0); S(wp00, qw, dw, 1, 0, 0); S(wm00, qw, dw, -1, 0, 0);
    S(w0p0, qw, dw, 0, 1, 0); S(w0m0, qw, dw, 0, -1, 0); S(w00p, qw, dw, 0, 0, 1);
    S(w00m, qw, dw, 0, 0, -1); S(wp0m, qw, dw, 1, 0, -1); S(w0pm, qw, dw, 0, 1, -1);
#undef S
    D1 fu, fv, fw;
    mom_rhs(g, n, k, fu, fv, fw);
    ou[e] = fu.d;
    ov[e] = fv.d;
    ow[e] = fw.d;
}

// ------------------------------------------------------------ boundary fills
// The forward fills set every ghost from interior values (periodic copies in
// x and y; at the walls u,v ghost = 2 bc - interior of the first/last plane,
// w ghost = 0; w's top wall row k = nz := 0; p ghost = periodic copy or the
// first/last plane).  The transpose adds each ghost's adjoint to the interior
// value it was made from and clears the ghosts.  Pass 1 (interior cells next
// to a face) gathers; pass 2 clears the ghost shell.
__device__ __forceinline__ int images(int i, int n, int out[3])
{
    int m = 0;
    out[m++] = i;
    if (i == n) out[m++] = 0;
    if (i == 1) out[m++] = n + 1;
    return m;
}

// kind 0: velocity (3 fields), kind 1: pressure (field 0 only).
template <int KIND>
__global__ void __launch_bounds__(256) k_fold(G g, double *__restrict__ f0,
                                              double *__restrict__ f1,
                                              double *__restrict__ f2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny * g.nz) return;
    const int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1,
              k = (int)(t / ((long long)g.nx * g.ny)) + 1;
    const bool face = i == 1 || i == g.nx || j == 1 || j == g.ny || k == 1 || k == g.nz;
    if (!face) return;
    int ia[3], ja[3];
    const int ni = images(i, g.nx, ia), nj = images(j, g.ny, ja);
    const long long sy = g.sy, sz = g.sz;
    const long long e = i + sy * j + sz * k;
    double *F[3] = {f0, f1, f2};
    const int nf = KIND == 0 ? 3 : 1;
    for (int f = 0; f < nf; ++f) {
        double *q = F[f];
        if (KIND == 0 && f == 2 && k == g.nz) {  // w's wall row is overwritten with 0
            q[e] = 0.0;
            continue;
        }
        double s = q[e];
        for (int a = 0; a < ni; ++a)
            for (int b = 0; b < nj; ++b)
                if (a != 0 || b != 0) s += q[ia[a] + sy * ja[b] + sz * k];
        // wall ghost planes: p copies (+1), u and v mirror (-1), w is set to 0
        const double wc = KIND == 1 ? 1.0 : (f == 2 ? 0.0 : -1.0);
        if (wc != 0.0) {
            for (int side = 0; side < 2; ++side) {
                if (side == 0 ? k != 1 : k != g.nz) continue;
                const int kg = side == 0 ? 0 : g.nz + 1;
                for (int a = 0; a < ni; ++a)
                    for (int b = 0; b < nj; ++b) s += wc * q[ia[a] + sy * ja[b] + sz * kg];
            }
        }
        q[e] = s;
    }
}

__global__ void __launch_bounds__(256) k_clear_ghosts(G g, double *__restrict__ f0,
                                                      double *__restrict__ f1,
                                                      double *__restrict__ f2, int nf)
{
    int i, j, k;
    if (!cell_of(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    if (interior(g, i, j, k)) return;
    const long long e = i + g.sy * j + g.sz * k;
    f0[e] = 0.0;
    if (nf > 1) {
        f1[e] = 0.0;
        f2[e] = 0.0;
    }
}

// ------------------------------------------------------------ projection
// correctVel: u -= (p_{i+1} - p_i) dtdxi, v -= (p_{j+1} - p_j) dtdyi,
// w -= (p_{k+1} - p_k) dt dzc_inv(k) on the interior.  Transpose: for every
// p cell c (interior or halo), lp(c) = sum of the coefficients times the
// adjoints of the interior velocities that read it.
__global__ void __launch_bounds__(256) k_correct_adj(G g, const double *__restrict__ lu,
                                                     const double *__restrict__ lv,
                                                     const double *__restrict__ lw,
                                                     double *__restrict__ lp, double dtdxi,
                                                     double dtdyi)
{
    int i, j, k;
    if (!cell_of(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long long sy = g.sy, sz = g.sz;
    const long long e = i + sy * j + sz * k;
    double s = 0.0;
    if (interior(g, i, j, k)) {
        s += dtdxi * lu[e] + dtdyi * lv[e] + g.dt * g.dzc_inv[k] * lw[e];
    }
    if (interior(g, i - 1, j, k)) s -= dtdxi * lu[e - 1];
    if (interior(g, i, j - 1, k)) s -= dtdyi * lv[e - sy];
    if (interior(g, i, j, k - 1)) s -= g.dt * g.dzc_inv[k - 1] * lw[e - sz];
    lp[e] = s;
}

// calcRHS: rhs = (u_i - u_{i-1}) dtidxi + (v_j - v_{j-1}) dtidyi
//              + (w_k - w_{k-1}) dti dzf_inv(k) on the interior.
// Transpose, accumulated: l_u(c) += dtidxi (lr(c) - lr(c+x)) etc., over every
// velocity cell c, with lr read on the interior only.
__global__ void __launch_bounds__(256) k_rhs_adj(G g, const double *__restrict__ lr,
                                                 double *__restrict__ lu, double *__restrict__ lv,
                                                 double *__restrict__ lw, double dti,
                                                 double dtidxi, double dtidyi)
{
    int i, j, k;
    if (!cell_of(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const long long sy = g.sy, sz = g.sz;
    const long long e = i + sy * j + sz * k;
    const double rc = interior(g, i, j, k) ? lr[e] : 0.0;
    const double rx = interior(g, i + 1, j, k) ? lr[e + 1] : 0.0;
    const double ry = interior(g, i, j + 1, k) ? lr[e + sy] : 0.0;
    const double rz = interior(g, i, j, k + 1) ? lr[e + sz] : 0.0;
    lu[e] += dtidxi * (rc - rx);
    lv[e] += dtidyi * (rc - ry);
    double wz = 0.0;
    if (k >= 1 && k <= g.nz) wz += dti * g.dzf_inv[k] * rc;
    if (k + 1 >= 1 && k + 1 <= g.nz) wz -= dti * g.dzf_inv[k + 1] * rz;
    lw[e] += wz;
}

// ------------------------------------------------------------ forcing
// forceVel: q += vel_force - sum_c q_c dzflzi(k_c) / (nx ny).  Transpose:
// l(c) -= dzflzi(k_c) / (nx ny) * sum_interior l.  Pass 1: per-(j,k)-row
// sums; the final sum and the shift in pass 2.
__global__ void __launch_bounds__(256) k_rowsum(G g, const double *__restrict__ q,
                                                double *__restrict__ part)
{
    __shared__ double sh[256];
    const int j = blockIdx.x % g.ny + 1, k = blockIdx.x / g.ny + 1;
    const long long base = g.sy * j + g.sz * k;
    double s = 0.0;
    for (int i = threadIdx.x + 1; i <= g.nx; i += blockDim.x) s += q[base + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(1024) k_sum(int n, const double *__restrict__ part,
                                              double *__restrict__ out, double scale)
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
    if (threadIdx.x == 0) out[0] = sh[0] * scale;
}

__global__ void __launch_bounds__(256) k_force_adj(G g, double *__restrict__ q,
                                                   const double *__restrict__ total)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny * g.nz) return;
    const int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1,
              k = (int)(t / ((long long)g.nx * g.ny)) + 1;
    const long long e = i + g.sy * j + g.sz * k;
    q[e] -= g.dzflzi[k] * total[0];
}

// ------------------------------------------------------------ wall shear
// sum over (i,j) of q(i,j,kw) for the bottom (kw = 1) and top (kw = nz)
// interior planes: per-row partials part[w * ny + j - 1].
__global__ void __launch_bounds__(256) k_wall_rows(G g, const double *__restrict__ q,
                                                   double *__restrict__ part)
{
    __shared__ double sh[256];
    const int w = blockIdx.x / g.ny, j = blockIdx.x % g.ny + 1;
    const int k = w == 0 ? 1 : g.nz;
    const long long base = g.sy * j + g.sz * k;
    double s = 0.0;
    for (int i = threadIdx.x + 1; i <= g.nx; i += blockDim.x) s += q[base + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// l_u += c_w on the interior planes k = 1 (c_bot) and k = nz (c_top).
__global__ void __launch_bounds__(256) k_wall_seed(G g, double *__restrict__ lu, double c_bot,
                                                   double c_top)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2LL * g.nx * g.ny) return;
    const int w = (int)(t / ((long long)g.nx * g.ny));
    const long long r = t % ((long long)g.nx * g.ny);
    const int i = (int)(r % g.nx) + 1, j = (int)(r / g.nx) + 1;
    const int k = w == 0 ? 1 : g.nz;
    if (g.nz == 1 && w == 1) {  // one plane: both walls are that plane
        lu[i + g.sy * j + g.sz * k] += c_top;
        return;
    }
    lu[i + g.sy * j + g.sz * k] += (w == 0 ? c_bot : c_top);
}

// y += a x over every cell of the haloed arrays (three fields).
__global__ void __launch_bounds__(256) k_axpy3(long long n, double a, const double *__restrict__ x0,
                                               const double *__restrict__ x1,
                                               const double *__restrict__ x2,
                                               double *__restrict__ y0, double *__restrict__ y1,
                                               double *__restrict__ y2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    y0[t] += a * x0[t];
    y1[t] += a * x1[t];
    y2[t] += a * x2[t];
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -3; }
inline long long ncell(const G &g) { return (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2); }
inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int pll_launch_jtv(const pll_grid_dev *gd, const double *const q[3],
                              const double *const l[3], double *const o[3], double alpha,
                              double a0, double beta, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_jtv, dim3(nblk(ncell(g))), dim3(256), 0, s, g, q[0], q[1], q[2], l[0],
                       l[1], l[2], o[0], o[1], o[2], alpha, a0, beta);
    return ok();
}

extern "C" int pll_launch_jvp(const pll_grid_dev *gd, const double *const q[3],
                              const double *const dq[3], double *const o[3], hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_jvp, dim3(nblk(ncell(g))), dim3(256), 0, s, g, q[0], q[1], q[2], dq[0],
                       dq[1], dq[2], o[0], o[1], o[2]);
    return ok();
}

extern "C" int pll_launch_bc_vel_adj(const pll_grid_dev *gd, double *lu, double *lv, double *lw,
                                     hipStream_t s)
{
    const G g = make_g(gd);
    const long long n = (long long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_fold<0>, dim3(nblk(n)), dim3(256), 0, s, g, lu, lv, lw);
    if (ok()) return -3;
    hipLaunchKernelGGL(k_clear_ghosts, dim3(nblk(ncell(g))), dim3(256), 0, s, g, lu, lv, lw, 3);
    return ok();
}

extern "C" int pll_launch_bc_p_adj(const pll_grid_dev *gd, double *lp, hipStream_t s)
{
    const G g = make_g(gd);
    const long long n = (long long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_fold<1>, dim3(nblk(n)), dim3(256), 0, s, g, lp, lp, lp);
    if (ok()) return -3;
    hipLaunchKernelGGL(k_clear_ghosts, dim3(nblk(ncell(g))), dim3(256), 0, s, g, lp, lp, lp, 1);
    return ok();
}

extern "C" int pll_launch_correct_adj(const pll_grid_dev *gd, const double *lu, const double *lv,
                                      const double *lw, double *lp, hipStream_t s)
{
    const G g = make_g(gd);
    const double dtdxi = g.dt * g.dx_inv, dtdyi = g.dt * g.dy_inv;
    hipLaunchKernelGGL(k_correct_adj, dim3(nblk(ncell(g))), dim3(256), 0, s, g, lu, lv, lw, lp,
                       dtdxi, dtdyi);
    return ok();
}

extern "C" int pll_launch_rhs_adj(const pll_grid_dev *gd, const double *lr, double *lu,
                                  double *lv, double *lw, hipStream_t s)
{
    const G g = make_g(gd);
    const double dti = 1.0 / g.dt;
    hipLaunchKernelGGL(k_rhs_adj, dim3(nblk(ncell(g))), dim3(256), 0, s, g, lr, lu, lv, lw, dti,
                       dti * g.dx_inv, dti * g.dy_inv);
    return ok();
}

extern "C" int pll_launch_force_adj(const pll_grid_dev *gd, double *q, double *part,
                                    double *total, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_rowsum, dim3(g.ny * g.nz), dim3(256), 0, s, g, q, part);
    if (ok()) return -3;
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, g.ny * g.nz, part, total,
                       1.0 / ((double)g.nx * g.ny));
    if (ok()) return -3;
    const long long n = (long long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_force_adj, dim3(nblk(n)), dim3(256), 0, s, g, q, total);
    return ok();
}

extern "C" int pll_launch_wall_sum(const pll_grid_dev *gd, const double *u, double *part,
                                   double *out, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_wall_rows, dim3(2 * g.ny), dim3(256), 0, s, g, u, part);
    if (ok()) return -3;
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, g.ny, part, out, 1.0);
    if (ok()) return -3;
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, g.ny, part + g.ny, out + 1, 1.0);
    return ok();
}

extern "C" int pll_launch_wall_seed(const pll_grid_dev *gd, double *lu, double c_bot,
                                    double c_top, hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_wall_seed, dim3(nblk(2LL * g.nx * g.ny)), dim3(256), 0, s, g, lu, c_bot,
                       c_top);
    return ok();
}

extern "C" int pll_launch_axpy3(const pll_grid_dev *gd, double a, const double *const x[3],
                                double *const y[3], hipStream_t s)
{
    const G g = make_g(gd);
    const long long n = ncell(g);
    hipLaunchKernelGGL(k_axpy3, dim3(nblk(n)), dim3(256), 0, s, n, a, x[0], x[1], x[2], y[0], y[1],
                       y[2]);
    return ok();
}
