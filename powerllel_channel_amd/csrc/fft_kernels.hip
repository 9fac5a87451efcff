// FFT stages of the pressure Poisson solve for gfx950.
//
// Replaces the FFTW r2r plans of src/src_c/mod_fft_fftw.inc:100-160 as used
// by execute_poisson_solver (src/src_c/mod_poisson_solver.c:484-494 x-R2HC,
// :510-514 y-R2HC, :568-571 y-HC2R, :587-599 x-HC2R + normalise + copy).
//
// Design: every transform is a complex Stockham FFT of length L done by T
// threads that hold E points each in registers (L = T*E).  Radix stages read
// their inputs from registers (stage 0) or LDS and write their outputs to
// LDS (the last stage keeps them in registers, already in natural order).
// A real row of nx points is transformed as the complex row
// z_m = x_{2m} + i x_{2m+1} of length L = nx/2 plus one LDS exchange that
// separates the even/odd spectra; the spectrum is stored as L complex slots,
// slot 0 holding the two real values (X_0, X_{nx/2}).  The y transform is a
// complex FFT down columns of that array; column 0 (two real spectra packed
// in one complex column) is split into its two Hermitian halves.
//
// All tables (twiddles) come from the host, computed in long double.
#include <hip/hip_runtime.h>

#include "pll_internal.h"

namespace {

typedef double2 cd;

__device__ __forceinline__ cd cadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cscale(cd a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ cd cconj(cd a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ cd cmul(cd a, cd b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// DIR * i * a
template <int DIR>
__device__ __forceinline__ cd rot(cd a)
{
    return DIR < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// Two doubles at an 8-byte (not 16-byte) aligned address: the interior of a
// haloed row starts one double past the row start, so a pair (x_2m, x_2m+1)
// is moved as one 16-byte access that is only 8-byte aligned.
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));

constexpr double kS3 = 0.86602540378443864676372317075294;  // sqrt(3)/2
constexpr double kH = 0.70710678118654752440084436210485;   // sqrt(2)/2

// In-register DFTs of size R with exp(DIR * 2 pi i / R) as the root.
template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
    static __device__ __forceinline__ void run(cd *x)
    {
        cd a = x[0];
        x[0] = cadd(a, x[1]);
        x[1] = csub(a, x[1]);
    }
};

template <int DIR>
struct Dft<3, DIR> {
    static __device__ __forceinline__ void run(cd *x)
    {
        cd s = cadd(x[1], x[2]);
        cd d = csub(x[1], x[2]);
        cd m = csub(x[0], cscale(s, 0.5));
        cd r = rot<DIR>(cscale(d, kS3));
        x[0] = cadd(x[0], s);
        x[1] = cadd(m, r);
        x[2] = csub(m, r);
    }
};

template <int DIR>
struct Dft<4, DIR> {
    static __device__ __forceinline__ void run(cd *x)
    {
        cd t0 = cadd(x[0], x[2]);
        cd t1 = csub(x[0], x[2]);
        cd t2 = cadd(x[1], x[3]);
        cd t3 = rot<DIR>(csub(x[1], x[3]));
        x[0] = cadd(t0, t2);
        x[2] = csub(t0, t2);
        x[1] = cadd(t1, t3);
        x[3] = csub(t1, t3);
    }
};

template <int DIR>
struct Dft<6, DIR> {
    static __device__ __forceinline__ void run(cd *x)
    {
        cd a[3] = {x[0], x[2], x[4]};
        cd b[3] = {x[1], x[3], x[5]};
        Dft<3, DIR>::run(a);
        Dft<3, DIR>::run(b);
        // W6^1 b1 = b1/2 + S3*rot(b1);  W6^2 b2 = -b2/2 + S3*rot(b2)
        cd b1 = cadd(cscale(b[1], 0.5), cscale(rot<DIR>(b[1]), kS3));
        cd b2 = csub(cscale(rot<DIR>(b[2]), kS3), cscale(b[2], 0.5));
        x[0] = cadd(a[0], b[0]);
        x[3] = csub(a[0], b[0]);
        x[1] = cadd(a[1], b1);
        x[4] = csub(a[1], b1);
        x[2] = cadd(a[2], b2);
        x[5] = csub(a[2], b2);
    }
};

template <int DIR>
struct Dft<8, DIR> {
    static __device__ __forceinline__ void run(cd *x)
    {
        cd e[4] = {x[0], x[2], x[4], x[6]};
        cd o[4] = {x[1], x[3], x[5], x[7]};
        Dft<4, DIR>::run(e);
        Dft<4, DIR>::run(o);
        cd o1 = cscale(cadd(o[1], rot<DIR>(o[1])), kH);
        cd o2 = rot<DIR>(o[2]);
        cd o3 = cscale(csub(rot<DIR>(o[3]), o[3]), kH);
        x[0] = cadd(e[0], o[0]);
        x[4] = csub(e[0], o[0]);
        x[1] = cadd(e[1], o1);
        x[5] = csub(e[1], o1);
        x[2] = cadd(e[2], o2);
        x[6] = csub(e[2], o2);
        x[3] = cadd(e[3], o3);
        x[7] = csub(e[3], o3);
    }
};

// One Stockham radix-R stage.  Thread t holds elements t + e*T (e < E).
// Butterfly b = t + q*T (q < E/R) takes elements b + r*L/R = slot q + r*E/R,
// twiddles them by w^(r*k) (k = b mod Ns, w = exp(DIR 2 pi i/(Ns R))) and
// writes output r to position (b - k)*R + k + r*Ns.  In the last stage that
// position is element t + (q + r*E/R)*T, i.e. the same register slot.
template <int R, int E, int DIR>
__device__ __forceinline__ void stage(cd (&v)[E], int t, int T, int Ns, int twstep,
                                      const cd *__restrict__ tw, cd *lds, bool last)
{
    if constexpr (E % R == 0) {
        constexpr int Q = E / R;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int b = t + q * T;
            const int k = b % Ns;
            cd x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = v[q + r * Q];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cd w = tw[r * k * twstep];
                    if (DIR > 0) w.y = -w.y;
                    x[r] = cmul(x[r], w);
                }
            }
            Dft<R, DIR>::run(x);
            if (last) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[q + r * Q] = x[r];
            } else {
                const int base = (b - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) lds[base + r * Ns] = x[r];
            }
        }
    }
}

// Full transform.  On entry v holds elements t + e*T of the input; on exit
// it holds elements t + e*T of the output.  `lds` is this transform's LDS
// segment (>= L points).  twscale = (table length) / L.  Every thread of the
// workgroup must call it (it contains barriers).
template <int E, int DIR>
__device__ __forceinline__ void fft_run(cd (&v)[E], int t, const pll_fftplan &P,
                                        const cd *__restrict__ tw, int twscale, cd *lds)
{
    const int T = P.T, L = P.L;
    int Ns = 1;
    for (int s = 0; s < P.nst; ++s) {
        const int R = P.rad[s];
        const bool last = (s == P.nst - 1);
        if (s > 0) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[t + e * T];
            __syncthreads();
        }
        const int twstep = twscale * (L / (Ns * R));
        switch (R) {
            case 2: stage<2, E, DIR>(v, t, T, Ns, twstep, tw, lds, last); break;
            case 3: stage<3, E, DIR>(v, t, T, Ns, twstep, tw, lds, last); break;
            case 4: stage<4, E, DIR>(v, t, T, Ns, twstep, tw, lds, last); break;
            case 6: stage<6, E, DIR>(v, t, T, Ns, twstep, tw, lds, last); break;
            case 8: stage<8, E, DIR>(v, t, T, Ns, twstep, tw, lds, last); break;
            default: break;
        }
        Ns *= R;
    }
}

// Compile-time form of the same stages: radices, Ns and T are constants, so
// the butterfly indices (b mod Ns, output positions, twiddle offsets) fold
// into shifts and small constants and nothing is spilled between radices.
template <int R, int E, int DIR, int T, int Ns, int TWSTEP, bool LAST>
__device__ __forceinline__ void stage_ct(cd (&v)[E], int t, const cd *__restrict__ tw, cd *lds)
{
    constexpr int Q = E / R;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = t + q * T;
        const int k = b % Ns;
        cd x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = v[q + r * Q];
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cd w = tw[r * k * TWSTEP];
                if (DIR > 0) w.y = -w.y;
                x[r] = cmul(x[r], w);
            }
        }
        Dft<R, DIR>::run(x);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q + r * Q] = x[r];
        } else {
            const int base = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[base + r * Ns] = x[r];
        }
    }
}

template <int L, int E, int DIR, int TWSCALE, int Ns, int R, int... Rest>
__device__ __forceinline__ void stages_ct(cd (&v)[E], int t, const cd *__restrict__ tw, cd *lds)
{
    static_assert(E % R == 0, "radix must divide E");
    constexpr int T = L / E;
    constexpr bool LAST = sizeof...(Rest) == 0;
    constexpr int TWSTEP = TWSCALE * (L / (Ns * R));
    stage_ct<R, E, DIR, T, Ns, TWSTEP, LAST>(v, t, tw, lds);
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = lds[t + e * T];
        __syncthreads();
        stages_ct<L, E, DIR, TWSCALE, Ns * R, Rest...>(v, t, tw, lds);
    }
}

// FFT policies used by the kernels: a compile-time plan, or the runtime plan.
template <int L_, int E_, int... Rs>
struct CtFFT {
    static constexpr int E = E_;
    static __device__ __forceinline__ int threads(const pll_fftplan &) { return L_ / E_; }
    static __device__ __forceinline__ int length(const pll_fftplan &) { return L_; }
    template <int DIR, int TWSCALE>
    static __device__ __forceinline__ void run(cd (&v)[E_], int t, const pll_fftplan &,
                                               const cd *__restrict__ tw, cd *lds)
    {
        stages_ct<L_, E_, DIR, TWSCALE, 1, Rs...>(v, t, tw, lds);
    }
};

template <int E_>
struct RtFFT {
    static constexpr int E = E_;
    static __device__ __forceinline__ int threads(const pll_fftplan &P) { return P.T; }
    static __device__ __forceinline__ int length(const pll_fftplan &P) { return P.L; }
    template <int DIR, int TWSCALE>
    static __device__ __forceinline__ void run(cd (&v)[E_], int t, const pll_fftplan &P,
                                               const cd *__restrict__ tw, cd *lds)
    {
        fft_run<E_, DIR>(v, t, P, tw, TWSCALE, lds);
    }
};

struct XArgs {
    pll_fftplan P;
    const cd *tw;       // [nx] exp(-2 pi i t/nx)
    long long nrows;    // ny*nz
    int ny, rows_per_wg, pitch;
    long long off1, off2;  // haloed strides: (nx+2)(ny+2), nx+2
    double norm;
};

struct YArgs {
    pll_fftplan P;
    const cd *tw;       // [ny] exp(-2 pi i t/ny)
    int L, ny, cols, ntiles, pitch;
};

template <int E>
constexpr int max_threads()
{
    return E >= 12 ? 256 : 512;
}
template <class F>
constexpr int max_threads_f()
{
    return max_threads<F::E>();
}

// x forward: haloed real rows -> packed half spectrum rows of W.
template <class F>
__global__ void __launch_bounds__(max_threads_f<F>())
k_xfwd(const double *__restrict__ p, cd *__restrict__ W, XArgs a)
{
    extern __shared__ cd lds[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), L = F::length(a.P);
    const int rl = threadIdx.x / T, t = threadIdx.x - rl * T;
    const long long row = (long long)blockIdx.x * a.rows_per_wg + rl;
    const bool valid = row < a.nrows;
    cd *ml = lds + rl * a.pitch;
    const long long j = valid ? row % a.ny : 0, k = valid ? row / a.ny : 0;
    const double *src = p + (k + 1) * a.off1 + (j + 1) * a.off2 + 1;
    cd v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = t + e * T;
        if (valid) {
            const d2u pr = *reinterpret_cast<const d2u *>(src + 2 * m);
            v[e] = make_double2(pr.x, pr.y);
        } else {
            v[e] = make_double2(0.0, 0.0);
        }
    }
    F::template run<-1, 2>(v, t, a.P, a.tw, ml);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) ml[t + e * T] = v[e];
    __syncthreads();
    cd *dst = W + row * L;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = t + e * T;
        const cd Z = v[e];
        cd X;
        if (m == 0) {
            X = make_double2(Z.x + Z.y, Z.x - Z.y);  // (X_0, X_L)
        } else {
            const cd Zc = cconj(ml[L - m]);
            const cd Ev = cscale(cadd(Z, Zc), 0.5);
            const cd D = csub(Z, Zc);
            const cd O = make_double2(0.5 * D.y, -0.5 * D.x);  // -i D / 2
            X = cadd(Ev, cmul(a.tw[m], O));
        }
        if (valid) dst[m] = X;
    }
}

// x backward: packed half spectrum rows -> haloed real rows * norm.
template <class F>
__global__ void __launch_bounds__(max_threads_f<F>())
k_xbwd(const cd *__restrict__ W, double *__restrict__ p, XArgs a)
{
    extern __shared__ cd lds[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), L = F::length(a.P);
    const int rl = threadIdx.x / T, t = threadIdx.x - rl * T;
    const long long row = (long long)blockIdx.x * a.rows_per_wg + rl;
    const bool valid = row < a.nrows;
    cd *ml = lds + rl * a.pitch;
    const long long j = valid ? row % a.ny : 0, k = valid ? row / a.ny : 0;
    const cd *src = W + row * L;
    cd v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = t + e * T;
        v[e] = valid ? src[m] : make_double2(0.0, 0.0);
        ml[m] = v[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = t + e * T;
        const cd X = v[e];
        if (m == 0) {
            v[e] = make_double2(X.x + X.y, X.x - X.y);
        } else {
            const cd Xc = cconj(ml[L - m]);
            const cd S = cadd(X, Xc);
            const cd D = csub(X, Xc);
            const cd Tt = cmul(cconj(a.tw[m]), D);
            v[e] = make_double2(S.x - Tt.y, S.y + Tt.x);  // S + i Tt
        }
    }
    __syncthreads();
    F::template run<+1, 2>(v, t, a.P, a.tw, ml);
    if (valid) {
        double *dst = p + (k + 1) * a.off1 + (j + 1) * a.off2 + 1;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            d2u pr;
            pr.x = v[e].x * a.norm;
            pr.y = v[e].y * a.norm;
            *reinterpret_cast<d2u *>(dst + 2 * m) = pr;
        }
    }
}

// y forward, in place on W: tile of `cols` adjacent columns of one k-plane.
template <class F>
__global__ void __launch_bounds__(max_threads_f<F>())
k_yfwd(cd *__restrict__ W, YArgs a)
{
    extern __shared__ cd lds[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), ny = F::length(a.P), L = a.L;
    const int tile = blockIdx.x % a.ntiles;
    const long long k = blockIdx.x / a.ntiles;
    const int c = threadIdx.x % a.cols, t = threadIdx.x / a.cols;
    const int mx = tile * a.cols + c;
    const bool valid = mx < L;
    cd *col = W + k * (long long)ny * L + (valid ? mx : 0);
    cd *ml = lds + c * a.pitch;
    cd v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int my = t + e * T;
        v[e] = valid ? col[(long long)my * L] : make_double2(0.0, 0.0);
    }
    F::template run<-1, 1>(v, t, a.P, a.tw, ml);
    if (tile == 0) {
        // column 0 holds DFT_y(X_0 + i X_L): split into A = DFT_y(X_0) and
        // B = DFT_y(X_L), keeping one half of each Hermitian sequence.
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) ml[t + e * T] = v[e];
        __syncthreads();
        if (c == 0) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int my = t + e * T;
                if (my == 0 || 2 * my == ny) continue;  // (A, B) = (Re C, Im C)
                if (2 * my < ny) {
                    v[e] = cscale(cadd(v[e], cconj(ml[ny - my])), 0.5);  // A_my
                } else {
                    const cd D = csub(ml[ny - my], cconj(v[e]));
                    v[e] = make_double2(0.5 * D.y, -0.5 * D.x);  // B_{ny-my}
                }
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; ++e) col[(long long)(t + e * T) * L] = v[e];
    }
}

template <class F>
__global__ void __launch_bounds__(max_threads_f<F>())
k_ybwd(cd *__restrict__ W, YArgs a)
{
    extern __shared__ cd lds[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), ny = F::length(a.P), L = a.L;
    const int tile = blockIdx.x % a.ntiles;
    const long long k = blockIdx.x / a.ntiles;
    const int c = threadIdx.x % a.cols, t = threadIdx.x / a.cols;
    const int mx = tile * a.cols + c;
    const bool valid = mx < L;
    cd *col = W + k * (long long)ny * L + (valid ? mx : 0);
    cd *ml = lds + c * a.pitch;
    cd v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int my = t + e * T;
        v[e] = valid ? col[(long long)my * L] : make_double2(0.0, 0.0);
    }
    if (tile == 0) {
        // rebuild C = A + i B over all my from the packed halves
#pragma unroll
        for (int e = 0; e < E; ++e) ml[t + e * T] = v[e];
        __syncthreads();
        if (c == 0) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int my = t + e * T;
                if (my == 0 || 2 * my == ny) continue;
                if (2 * my < ny) {
                    const cd B = ml[ny - my];
                    v[e] = make_double2(v[e].x - B.y, v[e].y + B.x);
                } else {
                    const cd A = cconj(ml[ny - my]);
                    const cd B = cconj(v[e]);
                    v[e] = make_double2(A.x - B.y, A.y + B.x);
                }
            }
        }
        __syncthreads();
    }
    F::template run<+1, 1>(v, t, a.P, a.tw, ml);
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; ++e) col[(long long)(t + e * T) * L] = v[e];
    }
}

template <class F>
int launch_x(bool fwd, const pll_ppe_dev *d, const double *p_in, double *p_out, void *W,
             hipStream_t s)
{
    XArgs a;
    a.P = d->px;
    a.tw = (const cd *)d->tw_x;
    a.nrows = (long long)d->ny * d->nz;
    a.ny = d->ny;
    a.rows_per_wg = d->x_rows_per_wg;
    a.pitch = d->px.L + 1;
    a.off1 = (long long)(d->nx + 2) * (d->ny + 2);
    a.off2 = d->nx + 2;
    a.norm = d->normfactor;
    const int threads = a.rows_per_wg * a.P.T;
    const long long blocks = (a.nrows + a.rows_per_wg - 1) / a.rows_per_wg;
    const size_t shm = (size_t)a.rows_per_wg * a.pitch * sizeof(cd);
    if (threads > max_threads_f<F>() || shm > 65536 || blocks > 0x7fffffffLL) return -1;
    if (fwd)
        hipLaunchKernelGGL(k_xfwd<F>, dim3((unsigned)blocks), dim3(threads), shm, s, p_in,
                           (cd *)W, a);
    else
        hipLaunchKernelGGL(k_xbwd<F>, dim3((unsigned)blocks), dim3(threads), shm, s,
                           (const cd *)W, p_out, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <class F>
int launch_y(bool fwd, const pll_ppe_dev *d, void *W, hipStream_t s)
{
    YArgs a;
    a.P = d->py;
    a.tw = (const cd *)d->tw_y;
    a.L = d->L;
    a.ny = d->ny;
    a.cols = d->y_cols_per_wg;
    a.ntiles = (d->L + a.cols - 1) / a.cols;
    a.pitch = d->ny + 1;
    const int threads = a.cols * a.P.T;
    const long long blocks = (long long)a.ntiles * d->nz;
    const size_t shm = (size_t)a.cols * a.pitch * sizeof(cd);
    if (threads > max_threads_f<F>() || shm > 65536 || blocks > 0x7fffffffLL) return -1;
    if (fwd)
        hipLaunchKernelGGL(k_yfwd<F>, dim3((unsigned)blocks), dim3(threads), shm, s, (cd *)W, a);
    else
        hipLaunchKernelGGL(k_ybwd<F>, dim3((unsigned)blocks), dim3(threads), shm, s, (cd *)W, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// Compile-time plans for the transform lengths of the production grids
// (x: L = nx/2, y: L = ny).  Every entry must match what ct_plan() reports.
#define PLL_CT_PLANS(X)                 \
    X(32, 8, 8, 4)                      \
    X(48, 24, 8, 6)                     \
    X(64, 8, 8, 8)                      \
    X(96, 24, 8, 4, 3)                  \
    X(128, 8, 8, 8, 2)                  \
    X(192, 24, 8, 8, 3)                 \
    X(192, 12, 4, 4, 4, 3)              \
    X(256, 8, 8, 8, 4)                  \
    X(384, 24, 8, 8, 6)                 \
    X(384, 12, 4, 4, 4, 6)              \
    X(512, 8, 8, 8, 8)                  \
    X(768, 24, 8, 8, 4, 3)              \
    X(768, 12, 4, 4, 4, 4, 3)           \
    X(1024, 8, 8, 8, 8, 2)

template <typename G>
int dispatch_plan(const pll_fftplan &P, G &&g)
{
#define PLL_CT_CASE(L_, E_, ...) \
    if (P.L == L_ && P.E == E_) return g(CtFFT<L_, E_, __VA_ARGS__>{});
    PLL_CT_PLANS(PLL_CT_CASE)
#undef PLL_CT_CASE
    switch (P.E) {
        case 2: return g(RtFFT<2>{});
        case 3: return g(RtFFT<3>{});
        case 4: return g(RtFFT<4>{});
        case 6: return g(RtFFT<6>{});
        case 8: return g(RtFFT<8>{});
        case 12: return g(RtFFT<12>{});
        case 24: return g(RtFFT<24>{});
        default: return -5;
    }
}

}  // namespace

extern "C" int pll_fft_supported_E(int E)
{
    return E == 2 || E == 3 || E == 4 || E == 6 || E == 8 || E == 12 || E == 24;
}

extern "C" int pll_fft_max_threads(int E) { return E >= 12 ? 256 : 512; }

extern "C" int pll_fft_ct_plan(int L, int prefer_E, pll_fftplan *p)
{
    const int plans[][8] = {
#define PLL_CT_ROW(L_, E_, ...) {L_, E_, __VA_ARGS__},
        PLL_CT_PLANS(PLL_CT_ROW)
#undef PLL_CT_ROW
    };
    int pick = -1;
    for (size_t i = 0; i < sizeof plans / sizeof plans[0]; i++) {
        if (plans[i][0] != L) continue;
        if (pick < 0 || plans[i][1] == prefer_E) pick = (int)i;
    }
    for (size_t i = (size_t)(pick < 0 ? 0 : pick); pick >= 0 && i == (size_t)pick; i++) {
        p->L = L;
        p->E = plans[i][1];
        p->T = L / p->E;
        int n = 0, prod = 1;
        for (int j = 2; j < 8 && plans[i][j] > 0; j++) {
            p->rad[n++] = plans[i][j];
            prod *= plans[i][j];
        }
        p->nst = n;
        return prod == L;
    }
    return 0;
}

extern "C" int pll_launch_xfwd(const pll_ppe_dev *d, const double *p, void *W, hipStream_t s)
{
    return dispatch_plan(d->px, [&](auto f) { return launch_x<decltype(f)>(true, d, p, nullptr, W, s); });
}

extern "C" int pll_launch_xbwd(const pll_ppe_dev *d, const void *W, double *p, hipStream_t s)
{
    return dispatch_plan(d->px, [&](auto f) {
        return launch_x<decltype(f)>(false, d, nullptr, p, const_cast<void *>(W), s);
    });
}

extern "C" int pll_launch_yfwd(const pll_ppe_dev *d, void *W, hipStream_t s)
{
    return dispatch_plan(d->py, [&](auto f) { return launch_y<decltype(f)>(true, d, W, s); });
}

extern "C" int pll_launch_ybwd(const pll_ppe_dev *d, void *W, hipStream_t s)
{
    return dispatch_plan(d->py, [&](auto f) { return launch_y<decltype(f)>(false, d, W, s); });
}
