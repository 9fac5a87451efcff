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
// Every kernel exists in double (the FP64 solve) and float (the FP32 solve
// of the refined form: FFT arithmetic and the spectral array in FP32, the
// real-space input and output in FP64).  Twiddle tables come from the host,
// computed in long double and rounded once to the working precision.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pll_internal.h"

namespace {

template <typename V>
struct Sc;
template <>
struct Sc<double2> {
    using T = double;
};
template <>
struct Sc<float2> {
    using T = float;
};

template <typename V>
__device__ __forceinline__ V mk(typename Sc<V>::T x, typename Sc<V>::T y)
{
    V r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
template <typename V>
__device__ __forceinline__ V cscale(V a, typename Sc<V>::T s) { return mk<V>(a.x * s, a.y * s); }
template <typename V>
__device__ __forceinline__ V cconj(V a) { return mk<V>(a.x, -a.y); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// The file is compiled without automatic contraction, so every
// instantiation of a transform (one per source/destination policy) rounds
// identically; the complex products use explicit fused multiply-adds.
template <typename V>
__device__ __forceinline__ V cmul(V a, V b)
{
    return mk<V>(fma_(a.x, b.x, -(a.y * b.y)), fma_(a.x, b.y, a.y * b.x));
}
// DIR * i * a
template <int DIR, typename V>
__device__ __forceinline__ V rot(V a)
{
    return DIR < 0 ? mk<V>(a.y, -a.x) : mk<V>(-a.y, a.x);
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
    template <typename V>
    static __device__ __forceinline__ void run(V *x)
    {
        V a = x[0];
        x[0] = cadd(a, x[1]);
        x[1] = csub(a, x[1]);
    }
};

template <int DIR>
struct Dft<3, DIR> {
    template <typename V>
    static __device__ __forceinline__ void run(V *x)
    {
        using S = typename Sc<V>::T;
        V s = cadd(x[1], x[2]);
        V d = csub(x[1], x[2]);
        V m = csub(x[0], cscale(s, S(0.5)));
        V r = rot<DIR>(cscale(d, S(kS3)));
        x[0] = cadd(x[0], s);
        x[1] = cadd(m, r);
        x[2] = csub(m, r);
    }
};

template <int DIR>
struct Dft<4, DIR> {
    template <typename V>
    static __device__ __forceinline__ void run(V *x)
    {
        V t0 = cadd(x[0], x[2]);
        V t1 = csub(x[0], x[2]);
        V t2 = cadd(x[1], x[3]);
        V t3 = rot<DIR>(csub(x[1], x[3]));
        x[0] = cadd(t0, t2);
        x[2] = csub(t0, t2);
        x[1] = cadd(t1, t3);
        x[3] = csub(t1, t3);
    }
};

template <int DIR>
struct Dft<6, DIR> {
    template <typename V>
    static __device__ __forceinline__ void run(V *x)
    {
        using S = typename Sc<V>::T;
        V a[3] = {x[0], x[2], x[4]};
        V b[3] = {x[1], x[3], x[5]};
        Dft<3, DIR>::run(a);
        Dft<3, DIR>::run(b);
        // W6^1 b1 = b1/2 + S3*rot(b1);  W6^2 b2 = -b2/2 + S3*rot(b2)
        V b1 = cadd(cscale(b[1], S(0.5)), cscale(rot<DIR>(b[1]), S(kS3)));
        V b2 = csub(cscale(rot<DIR>(b[2]), S(kS3)), cscale(b[2], S(0.5)));
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
    template <typename V>
    static __device__ __forceinline__ void run(V *x)
    {
        using S = typename Sc<V>::T;
        V e[4] = {x[0], x[2], x[4], x[6]};
        V o[4] = {x[1], x[3], x[5], x[7]};
        Dft<4, DIR>::run(e);
        Dft<4, DIR>::run(o);
        V o1 = cscale(cadd(o[1], rot<DIR>(o[1])), S(kH));
        V o2 = rot<DIR>(o[2]);
        V o3 = cscale(csub(rot<DIR>(o[3]), o[3]), S(kH));
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

// Odd prime radices R = 5, 7, 11, 13: cos(2 pi p/R) and sin(2 pi p/R) for
// p = 1..(R-1)/2.
template <int R>
struct OddRoots;
template <>
struct OddRoots<5> {
    static constexpr double c[2] = {0.30901699437494742410, -0.80901699437494742410};
    static constexpr double s[2] = {0.95105651629515357212, 0.58778525229247312917};
};
template <>
struct OddRoots<7> {
    static constexpr double c[3] = {0.62348980185873353053, -0.22252093395631440429,
                                    -0.90096886790241912624};
    static constexpr double s[3] = {0.78183148246802980871, 0.97492791218182360702,
                                    0.43388373911755812048};
};
template <>
struct OddRoots<11> {
    static constexpr double c[5] = {0.84125353283118116886, 0.41541501300188642553,
                                    -0.14231483827328514044, -0.65486073394528506406,
                                    -0.95949297361449738989};
    static constexpr double s[5] = {0.54064081745559758211, 0.90963199535451837141,
                                    0.98982144188093273238, 0.75574957435425828377,
                                    0.28173255684142969771};
};
template <>
struct OddRoots<13> {
    static constexpr double c[6] = {0.88545602565320989590, 0.56806474673115580251,
                                    0.12053668025532305335, -0.35460488704253562597,
                                    -0.74851074817110109863, -0.97094181742605202716};
    static constexpr double s[6] = {0.46472317204376854566, 0.82298386589365639458,
                                    0.99270887409805399280, 0.93501624268541482344,
                                    0.66312265824079520238, 0.23931566428755776715};
};

// Odd prime R: X_m = x_0 + sum_{j=1..H} (cos(2 pi jm/R) (x_j + x_{R-j})
//                        + DIR i sin(2 pi jm/R) (x_j - x_{R-j})), H = (R-1)/2,
// and X_{R-m} the same with the sine part negated.
template <int R, int DIR>
struct DftOdd {
    template <typename V>
    static __device__ __forceinline__ void run(V *x)
    {
        using S = typename Sc<V>::T;
        constexpr int H = (R - 1) / 2;
        V tp[H], dm[H];
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            tp[j - 1] = cadd(x[j], x[R - j]);
            dm[j - 1] = csub(x[j], x[R - j]);
        }
        const V x0 = x[0];
        V sum = x0;
#pragma unroll
        for (int j = 0; j < H; ++j) sum = cadd(sum, tp[j]);
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            V a = x0, b = mk<V>(S(0), S(0));
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const int p = (j * m) % R;
                const S cv = S(p <= H ? OddRoots<R>::c[p - 1] : OddRoots<R>::c[R - p - 1]);
                const S sv = S(p <= H ? OddRoots<R>::s[p - 1] : -OddRoots<R>::s[R - p - 1]);
                a = mk<V>(fma_(cv, tp[j - 1].x, a.x), fma_(cv, tp[j - 1].y, a.y));
                b = mk<V>(fma_(sv, dm[j - 1].x, b.x), fma_(sv, dm[j - 1].y, b.y));
            }
            const V rb = rot<DIR>(b);
            x[m] = cadd(a, rb);
            x[R - m] = csub(a, rb);
        }
        x[0] = sum;
    }
};
template <int DIR>
struct Dft<5, DIR> : DftOdd<5, DIR> {};
template <int DIR>
struct Dft<7, DIR> : DftOdd<7, DIR> {};
template <int DIR>
struct Dft<11, DIR> : DftOdd<11, DIR> {};
template <int DIR>
struct Dft<13, DIR> : DftOdd<13, DIR> {};

// One Stockham radix-R stage.  Thread t holds elements t + e*T (e < E).
// Butterfly b = t + q*T (q < E/R) takes elements b + r*L/R = slot q + r*E/R,
// twiddles them by w^(r*k) (k = b mod Ns, w = exp(DIR 2 pi i/(Ns R))) and
// writes output r to position (b - k)*R + k + r*Ns.  In the last stage that
// position is element t + (q + r*E/R)*T, i.e. the same register slot.
template <int R, int E, int DIR, typename V>
__device__ __forceinline__ void stage(V (&v)[E], int t, int T, int Ns, int twstep,
                                      const V *__restrict__ tw, V *lds, bool last)
{
    if constexpr (E % R == 0) {
        constexpr int Q = E / R;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int b = t + q * T;
            const int k = b % Ns;
            V x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = v[q + r * Q];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    V w = tw[r * k * twstep];
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

// The same stage when R does not divide E: the row is in LDS in natural
// element order; thread t takes butterflies b = t + qT < L/R (at most
// ceil(E/R) of them), reads their R inputs each, and after a barrier writes
// the outputs back to LDS at the Stockham positions.
template <int R, int E, int DIR, typename V>
__device__ __forceinline__ void stage_lds(int t, int T, int L, int Ns, int twstep,
                                          const V *__restrict__ tw, V *lds)
{
    constexpr int QM = (E + R - 1) / R;
    const int nb = L / R;
    V x[QM][R];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
        const int b = t + q * T;
        if (b < nb) {
            const int k = b % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) x[q][r] = lds[b + r * nb];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    V w = tw[r * k * twstep];
                    if (DIR > 0) w.y = -w.y;
                    x[q][r] = cmul(x[q][r], w);
                }
            }
            Dft<R, DIR>::run(x[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QM; ++q) {
        const int b = t + q * T;
        if (b < nb) {
            const int k = b % Ns;
            const int base = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[base + r * Ns] = x[q][r];
        }
    }
}

// Stage s of a runtime plan.  from_lds: the previous stage left the row in
// LDS (every stage but the first).  A radix dividing E runs from registers;
// any other radix runs from LDS, and as the last stage brings the result
// back into registers.
template <int R, int E, int DIR, typename V>
__device__ __forceinline__ void stage_rt(V (&v)[E], int t, int T, int L, int Ns, int twstep,
                                         const V *__restrict__ tw, V *lds, bool from_lds,
                                         bool last)
{
    if constexpr (E % R == 0) {
        if (from_lds) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[t + e * T];
            __syncthreads();
        }
        stage<R, E, DIR>(v, t, T, Ns, twstep, tw, lds, last);
    } else {
        if (!from_lds) {
#pragma unroll
            for (int e = 0; e < E; ++e) lds[t + e * T] = v[e];
        }
        __syncthreads();
        stage_lds<R, E, DIR>(t, T, L, Ns, twstep, tw, lds);
        if (last) {
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[t + e * T];
        }
    }
}

// Full transform with a runtime plan.  On entry v holds elements t + e*T of
// the input; on exit it holds elements t + e*T of the output.  `lds` is this
// transform's LDS segment (>= L points).  twscale = (table length) / L.
// Every thread of the workgroup must call it (it contains barriers).
template <int E, int DIR, typename V>
__device__ __forceinline__ void fft_run(V (&v)[E], int t, const pll_fftplan &P,
                                        const V *__restrict__ tw, int twscale, V *lds)
{
    const int T = P.T, L = P.L;
    int Ns = 1;
    for (int s = 0; s < P.nst; ++s) {
        const int R = P.rad[s];
        const bool last = (s == P.nst - 1), fl = s > 0;
        const int twstep = twscale * (L / (Ns * R));
        switch (R) {
            case 2: stage_rt<2, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 3: stage_rt<3, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 4: stage_rt<4, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 5: stage_rt<5, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 6: stage_rt<6, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 7: stage_rt<7, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 8: stage_rt<8, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 11: stage_rt<11, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            case 13: stage_rt<13, E, DIR>(v, t, T, L, Ns, twstep, tw, lds, fl, last); break;
            default: break;
        }
        Ns *= R;
    }
}

// Compile-time form of the same stages: radices, Ns and T are constants, so
// the butterfly indices (b mod Ns, output positions, twiddle offsets) fold
// into shifts and small constants and nothing is spilled between radices.
template <int R, int E, int DIR, int T, int Ns, int TWSTEP, bool LAST, typename V>
__device__ __forceinline__ void stage_ct(V (&v)[E], int t, const V *__restrict__ tw)
{
    constexpr int Q = E / R;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = t + q * T;
        const int k = b % Ns;
        V x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = v[q + r * Q];
        if constexpr (Ns > 1) {
            // w^r from one table entry w: powers by products of depth <= 3
            // (w^2 = w w, w^3 = w w^2, w^4 = w^2 w^2, ...), one load per butterfly
            V wp[R];
            wp[1] = tw[k * TWSTEP];
            if (DIR > 0) wp[1].y = -wp[1].y;
#pragma unroll
            for (int r = 2; r < R; ++r) wp[r] = cmul(wp[r / 2], wp[r - r / 2]);
#pragma unroll
            for (int r = 1; r < R; ++r) x[r] = cmul(x[r], wp[r]);
        }
        Dft<R, DIR>::run(x);
#pragma unroll
        for (int r = 0; r < R; ++r) v[q + r * Q] = x[r];
    }
}

// Between stages, output r of butterfly q (in v[q + r*Q]) moves to position
// (b - k) R + k + r Ns of the row (b = t + qT, k = b mod Ns), and element
// t + eT of the next stage comes back.  The exchange goes through LDS one
// real component at a time, so a row needs L scalars of LDS, not L complex
// values: twice the workgroups fit a CU.
template <typename V, int L, int E, int DIR, int TWSCALE, int Ns, int R, int... Rest>
__device__ __forceinline__ void stages_ct(V (&v)[E], int t, const V *__restrict__ tw,
                                          typename Sc<V>::T *ls)
{
    static_assert(E % R == 0, "radix must divide E");
    constexpr int T = L / E, Q = E / R;
    constexpr bool LAST = sizeof...(Rest) == 0;
    constexpr int TWSTEP = TWSCALE * (L / (Ns * R));
    stage_ct<R, E, DIR, T, Ns, TWSTEP, LAST>(v, t, tw);
    if constexpr (!LAST) {
        auto put = [&](bool imag) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int b = t + q * T;
                const int k = b % Ns;
                const int base = (b - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    ls[base + r * Ns] = imag ? v[q + r * Q].y : v[q + r * Q].x;
            }
        };
        put(false);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e].x = ls[t + e * T];
        __syncthreads();
        put(true);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e].y = ls[t + e * T];
        __syncthreads();
        stages_ct<V, L, E, DIR, TWSCALE, Ns * R, Rest...>(v, t, tw, ls);
    }
}

// FFT policies used by the kernels: a compile-time plan, or the runtime plan.
// kSplit: the plan exchanges through LDS one real component at a time (a row
// segment holds L+1 scalars); otherwise complex values (L+1 of them).
template <int L_, int E_, int... Rs>
struct CtFFT {
    static constexpr int E = E_;
    static constexpr int kT = L_ / E_;  // threads per row
    static constexpr bool kSplit = true;
    static __device__ __forceinline__ int threads(const pll_fftplan &) { return L_ / E_; }
    static __device__ __forceinline__ int length(const pll_fftplan &) { return L_; }
    template <int DIR, int TWSCALE, typename V>
    static __device__ __forceinline__ void run(V (&v)[E_], int t, const pll_fftplan &,
                                               const V *__restrict__ tw, void *lds)
    {
        stages_ct<V, L_, E_, DIR, TWSCALE, 1, Rs...>(v, t, tw,
                                                     reinterpret_cast<typename Sc<V>::T *>(lds));
    }
};

template <int E_>
struct RtFFT {
    static constexpr int E = E_;
    static constexpr int kT = 0;  // threads per row known at run time only
    static constexpr bool kSplit = false;
    static __device__ __forceinline__ int threads(const pll_fftplan &P) { return P.T; }
    static __device__ __forceinline__ int length(const pll_fftplan &P) { return P.L; }
    template <int DIR, int TWSCALE, typename V>
    static __device__ __forceinline__ void run(V (&v)[E_], int t, const pll_fftplan &P,
                                               const V *__restrict__ tw, void *lds)
    {
        fft_run<E_, DIR>(v, t, P, tw, TWSCALE, reinterpret_cast<V *>(lds));
    }
};

// Bytes of LDS per element of a row segment for plan F.
template <class F, typename V>
constexpr size_t lds_elem()
{
    return F::kSplit ? sizeof(typename Sc<V>::T) : sizeof(V);
}

// The x kernels' rows are T consecutive lanes of one wavefront when T is a
// compile-time power of two up to 64: their mirror can then go through lane
// shuffles instead of LDS (no barriers).
template <class F>
constexpr bool shfl_mirror()
{
    return F::kT > 0 && F::kT <= 64 && (F::kT & (F::kT - 1)) == 0;
}

// mv[e] := value at position n - m of the row (m = t + eT; position 0 for
// m = 0), where v[e] holds the value at m: the mirror the real-to-complex
// packing needs.  LDS form: every thread of the workgroup calls it
// (barriers); the caller synchronises before the row segment is written
// again.  SHFL form (rows of T consecutive lanes, shfl_mirror<F>()): for
// t >= 1 position n - m = (T - t) + (E - 1 - e) T is in lane T - t of the
// row, register E - 1 - e; for t = 0 it is (E - e) T, this lane's register
// (E - e) mod E.  No LDS, no barrier.
template <class F, typename V, int E, bool SHFL = false>
__device__ __forceinline__ void mirror(const V (&v)[E], int t, int T, int n, unsigned char *row,
                                       V (&mv)[E])
{
    using S = typename Sc<V>::T;
    if constexpr (SHFL) {
        static_assert(shfl_mirror<F>(), "rows of T consecutive lanes, T a power of two <= 64");
        constexpr int TT = F::kT;
        const int src = (int)(threadIdx.x & 63) - t + ((TT - t) & (TT - 1));
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const V a = v[E - 1 - e];
            const S ax = __shfl(a.x, src), ay = __shfl(a.y, src);
            const V b = v[(E - e) % E];
            mv[e] = t == 0 ? b : mk<V>(ax, ay);
        }
    } else if constexpr (F::kSplit) {
        S *ls = reinterpret_cast<S *>(row);
#pragma unroll
        for (int e = 0; e < E; ++e) ls[t + e * T] = v[e].x;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            mv[e].x = ls[m ? n - m : 0];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) ls[t + e * T] = v[e].y;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            mv[e].y = ls[m ? n - m : 0];
        }
    } else {
        V *lv = reinterpret_cast<V *>(row);
#pragma unroll
        for (int e = 0; e < E; ++e) lv[t + e * T] = v[e];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            mv[e] = lv[m ? n - m : 0];
        }
    }
}

// ---------------------------------------------------------------- x I/O
// Sources of the x forward transform: real row (j,k) of nx values.
// kLow: the source also gives the FP64 values (load_d), from which the x
// forward transform of the refined form projects the lowest modes in FP64.
struct SrcHalo64 {  // interior of a haloed FP64 array (the caller's p / rhs)
    static constexpr bool kLow = true;
    const double *p;
    long long off1, off2;
    __device__ __forceinline__ double2 load_d(long long, long long j, long long k, int m, int) const
    {
        const d2u pr = *reinterpret_cast<const d2u *>(p + (k + 1) * off1 + (j + 1) * off2 + 1 + 2 * m);
        return make_double2(pr.x, pr.y);
    }
    template <typename V>
    __device__ __forceinline__ V load(long long row, long long j, long long k, int m, int nx) const
    {
        const double2 pr = load_d(row, j, k, m, nx);
        return mk<V>((typename Sc<V>::T)pr.x, (typename Sc<V>::T)pr.y);
    }
};
struct SrcHalo64Copy {  // SrcHalo64, also copying the FP64 values to a dense array
    static constexpr bool kLow = true;
    const double *p;
    long long off1, off2;
    double *F;
    __device__ __forceinline__ double2 load_d(long long row, long long j, long long k, int m,
                                              int nx) const
    {
        const d2u pr = *reinterpret_cast<const d2u *>(p + (k + 1) * off1 + (j + 1) * off2 + 1 + 2 * m);
        *reinterpret_cast<double2 *>(F + row * nx + 2 * m) = make_double2(pr.x, pr.y);
        return make_double2(pr.x, pr.y);
    }
    template <typename V>
    __device__ __forceinline__ V load(long long row, long long j, long long k, int m, int nx) const
    {
        const double2 pr = load_d(row, j, k, m, nx);
        return mk<V>((typename Sc<V>::T)pr.x, (typename Sc<V>::T)pr.y);
    }
};
// Residual r = f - L_h p of a refinement sweep, formed in the load of the x
// transform: the arithmetic of k_resid (neighbour differences, row sums),
// rounded to the transform's precision.  f and p are each haloed or dense
// (element (i,j,k) at base + k s1 + j s2 + i).
struct SrcResid {
    static constexpr bool kLow = true;
    static constexpr bool kStencil = true;
    const double *f, *P;
    long long fb, f1, f2, pb, p1, p2;
    const double *a, *c, *d;
    double idx2, idy2;
    int ny, nz;
    template <typename V>
    __device__ __forceinline__ V load(long long row, long long j, long long k, int m, int nx) const
    {
        const double2 r = load_d(row, j, k, m, nx);
        return mk<V>((typename Sc<V>::T)(float)r.x, (typename Sc<V>::T)(float)r.y);
    }
    __device__ __forceinline__ double2 load_d(long long, long long j, long long k, int m,
                                              int nx) const
    {
        const int i0 = 2 * m, i1 = i0 + 1;
        const int im = i0 == 0 ? nx - 1 : i0 - 1, ip = i1 + 1 == nx ? 0 : i1 + 1;
        const long long jp = j + 1 == ny ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
        const double *__restrict__ row = P + pb + k * p1 + j * p2;
        const double *__restrict__ rjp = P + pb + k * p1 + jp * p2;
        const double *__restrict__ rjm = P + pb + k * p1 + jm * p2;
        // pairs (i0, i1) as one 16-byte access each (8-byte aligned in a
        // haloed array)
        auto pair = [](const double *q) { return *reinterpret_cast<const d2u *>(q); };
        const d2u cc = pair(row + i0), yp = pair(rjp + i0), ym = pair(rjm + i0);
        const double c0 = cc.x, c1 = cc.y;
        const double xm = row[im], xp = row[ip];
        const double yp0 = yp.x, yp1 = yp.y, ym0 = ym.x, ym1 = ym.y;
        double zm0 = 0.0, zm1 = 0.0, zp0 = 0.0, zp1 = 0.0;
        if (k > 0) {
            const d2u zm = pair(row + i0 - p1);
            zm0 = zm.x;
            zm1 = zm.y;
        }
        if (k + 1 < nz) {
            const d2u zp = pair(row + i0 + p1);
            zp0 = zp.x;
            zp1 = zp.y;
        }
        const double ak = a[k], ck = c[k], dk = d[k];
        const double *__restrict__ fr = f + fb + k * f1 + j * f2;
        const d2u ff = pair(fr + i0);
        const double f0 = ff.x, fv1 = ff.y;
        const double lx0 = ((c1 - c0) + (xm - c0)) * idx2;
        const double lx1 = ((xp - c1) + (c0 - c1)) * idx2;
        const double ly0 = ((yp0 - c0) + (ym0 - c0)) * idy2;
        const double ly1 = ((yp1 - c1) + (ym1 - c1)) * idy2;
        const double lz0 = ak * (zm0 - c0) + ck * (zp0 - c0) + dk * c0;
        const double lz1 = ak * (zm1 - c1) + ck * (zp1 - c1) + dk * c1;
        const double r0 = f0 - (lx0 + ly0 + lz0), r1 = fv1 - (lx1 + ly1 + lz1);
        return make_double2(r0, r1);
    }
};

struct SrcDense32 {  // dense FP32 rows (the residual of a refinement sweep)
    static constexpr bool kLow = false;
    const float *r;
    template <typename V>
    __device__ __forceinline__ V load(long long row, long long, long long, int m, int nx) const
    {
        const float2 pr = *reinterpret_cast<const float2 *>(r + row * nx + 2 * m);
        return mk<V>((typename Sc<V>::T)pr.x, (typename Sc<V>::T)pr.y);
    }
};

struct SrcDiv {  // PPE right-hand side formed from the velocity on the fly
    static constexpr bool kLow = false;
    pll_div_src q;
    long long off1, off2;
    template <typename V>
    __device__ __forceinline__ V load(long long, long long j, long long k, int m, int) const
    {
        // mod_calcRHS.f90:28-30, same operation order as the calcRHS kernel:
        //   rhs = (u - u_{i-1}) dtidxi + (v - v_{j-1}) dtidyi + (w - w_{k-1}) dti dzf_inv_k
#pragma clang fp contract(off)
        const long long e = (k + 1) * off1 + (j + 1) * off2 + 1 + 2 * m;
        const d2u uc = *reinterpret_cast<const d2u *>(q.u + e);
        const double um = q.u[e - 1];
        const d2u vc = *reinterpret_cast<const d2u *>(q.v + e);
        const d2u vs = *reinterpret_cast<const d2u *>(q.v + e - off2);
        const d2u wc = *reinterpret_cast<const d2u *>(q.w + e);
        const d2u wb = *reinterpret_cast<const d2u *>(q.w + e - off1);
        const double dzi = q.dzf_inv[k + 1];
        const double r0 = (uc.x - um) * q.dtidxi + (vc.x - vs.x) * q.dtidyi + (wc.x - wb.x) * q.dti * dzi;
        const double r1 = (uc.y - uc.x) * q.dtidxi + (vc.y - vs.y) * q.dtidyi + (wc.y - wb.y) * q.dti * dzi;
        return mk<V>((typename Sc<V>::T)r0, (typename Sc<V>::T)r1);
    }
};
struct SrcDivUnfilled {  // SrcDiv on a velocity whose boundary fill is not made
    static constexpr bool kLow = false;
    pll_div_src q;
    long long off1, off2;
    int nx, ny, nz;
    template <typename V>
    __device__ __forceinline__ V load(long long, long long j, long long k, int m, int) const
    {
        // the values updateBoundVel would have left: x and y periodic copies,
        // w = 0 on the wall rows (k = 0 below the first cell, k = nz on top)
#pragma clang fp contract(off)
        const long long e = (k + 1) * off1 + (j + 1) * off2 + 1 + 2 * m;
        const d2u uc = *reinterpret_cast<const d2u *>(q.u + e);
        const double um = q.u[m == 0 ? e - 1 + nx : e - 1];
        const d2u vc = *reinterpret_cast<const d2u *>(q.v + e);
        const d2u vs = *reinterpret_cast<const d2u *>(q.v + (j == 0 ? e + (long long)(ny - 1) * off2
                                                                     : e - off2));
        d2u wc = *reinterpret_cast<const d2u *>(q.w + e);
        if (k == nz - 1) wc.x = wc.y = 0.0;
        d2u wb;
        wb.x = wb.y = 0.0;
        if (k > 0) wb = *reinterpret_cast<const d2u *>(q.w + e - off1);
        const double dzi = q.dzf_inv[k + 1];
        const double r0 = (uc.x - um) * q.dtidxi + (vc.x - vs.x) * q.dtidyi + (wc.x - wb.x) * q.dti * dzi;
        const double r1 = (uc.y - uc.x) * q.dtidxi + (vc.y - vs.y) * q.dtidyi + (wc.y - wb.y) * q.dti * dzi;
        return mk<V>((typename Sc<V>::T)r0, (typename Sc<V>::T)r1);
    }
};

// Destinations of the x backward transform; (a, b) = normalised (x_2m, x_2m+1).
// kAligned: the destination takes 16-byte aligned pairs (x_{2m-1}, x_{2m})
// (row_pair) and single values (row_one) instead of (x_{2m}, x_{2m+1}) pairs,
// whose haloed slots start 8 bytes into a 16-byte unit
struct DstHalo64 {  // haloed FP64 array interior := value
    static constexpr bool kAligned = true;
    double *p;
    long long off1, off2;
    __device__ __forceinline__ void row_pair(long long j, long long k, int hi, double a,
                                             double b) const
    {
        d2u pr;
        pr.x = a;
        pr.y = b;
        *reinterpret_cast<d2u *>(p + (k + 1) * off1 + (j + 1) * off2 + hi) = pr;
    }
    __device__ __forceinline__ void row_one(long long j, long long k, int hi, double a) const
    {
        p[(k + 1) * off1 + (j + 1) * off2 + hi] = a;
    }
    __device__ __forceinline__ void store(long long, long long j, long long k, int m, int,
                                          double a, double b) const
    {
        d2u pr;
        pr.x = a;
        pr.y = b;
        *reinterpret_cast<d2u *>(p + (k + 1) * off1 + (j + 1) * off2 + 1 + 2 * m) = pr;
    }
};
struct DstHaloFill {  // haloed FP64 interior := value, plus the boundary fill made from it
    static constexpr bool kAligned = true;
    double *p;
    long long off1, off2;
    int nx, ny, nz;
    // updateBoundP at one rank: periodic images in x and y, zero-gradient
    // copies of the first/last plane into the ghost planes, corners included.
    // Every value of a row also goes to the row's images (the other j ghost
    // row, the ghost planes); x_0 and x_{nx-1} also to the row's x ghosts.
    template <class Fn>
    __device__ __forceinline__ void rows(long long j, long long k, Fn &&put) const
    {
        const int jj = (int)j + 1, kk = (int)k + 1;
        const int jo = jj == ny ? 0 : (jj == 1 ? ny + 1 : -1);
        put(kk * off1 + jj * off2);
        if (jo >= 0) put(kk * off1 + jo * off2);
        for (int side = 0; side < 2; ++side) {
            if (side == 0 ? kk != 1 : kk != nz) continue;
            const long long pg = (side == 0 ? 0 : nz + 1) * off1;
            put(pg + jj * off2);
            if (jo >= 0) put(pg + jo * off2);
        }
    }
    __device__ __forceinline__ void row_pair(long long j, long long k, int hi, double a,
                                             double b) const
    {
        d2u pr;
        pr.x = a;
        pr.y = b;
        rows(j, k, [&](long long rb) { *reinterpret_cast<d2u *>(p + rb + hi) = pr; });
    }
    __device__ __forceinline__ void row_one(long long j, long long k, int hi, double a) const
    {
        // x_0 (hi = 1) is also the ghost at nx + 1; x_{nx-1} (hi = nx) the ghost at 0
        rows(j, k, [&](long long rb) {
            p[rb + hi] = a;
            p[rb + (hi == 1 ? nx + 1 : 0)] = a;
        });
    }
};
struct DstDenseSet {  // dense FP64 := value
    static constexpr bool kAligned = false;
    double *P;
    __device__ __forceinline__ void store(long long row, long long, long long, int m, int nx,
                                          double a, double b) const
    {
        *reinterpret_cast<double2 *>(P + row * nx + 2 * m) = make_double2(a, b);
    }
};
struct DstDenseAdd {  // dense FP64 += value
    static constexpr bool kAligned = false;
    double *P;
    __device__ __forceinline__ void store(long long row, long long, long long, int m, int nx,
                                          double a, double b) const
    {
        double2 *q = reinterpret_cast<double2 *>(P + row * nx + 2 * m);
        const double2 o = *q;
        *q = make_double2(o.x + a, o.y + b);
    }
};
struct DstHaloFromDense {  // haloed FP64 interior := dense FP64 + value - eps
    static constexpr bool kAligned = false;
    double *p;
    long long off1, off2;
    const double *P;
    const double *eps;  // gauge constant (device scalar), may be null
    __device__ __forceinline__ void store(long long row, long long j, long long k, int m, int nx,
                                          double a, double b) const
    {
        const double2 o = *reinterpret_cast<const double2 *>(P + row * nx + 2 * m);
        const double e = eps ? *eps : 0.0;
        d2u pr;
        pr.x = (o.x - e) + a;
        pr.y = (o.y - e) + b;
        *reinterpret_cast<d2u *>(p + (k + 1) * off1 + (j + 1) * off2 + 1 + 2 * m) = pr;
    }
};

// Destinations that also report max |result| of the pair (store_n); the
// kernel adds max |value| and reduces both per workgroup.
__device__ __forceinline__ double nmax(double m, double x)
{
    return (x > m || x != x) ? x : m;  // NaN wins
}
struct DstHaloAddN {  // haloed FP64 interior := (interior - eps) + value
    static constexpr bool kAligned = false;
    double *p;
    long long off1, off2;
    const double *eps;  // gauge constant (device scalar), may be null
    __device__ __forceinline__ double store_n(long long, long long j, long long k, int m, int,
                                              double a, double b) const
    {
        d2u *q = reinterpret_cast<d2u *>(p + (k + 1) * off1 + (j + 1) * off2 + 1 + 2 * m);
        const d2u o = *q;
        const double e = eps ? *eps : 0.0;
        d2u pr;
        pr.x = (o.x - e) + a;
        pr.y = (o.y - e) + b;
        *q = pr;
        return nmax(fabs(pr.x), fabs(pr.y));
    }
};

template <typename V>
struct XArgs {
    pll_fftplan P;
    const V *tw;        // [nx] exp(-2 pi i t/nx)
    long long row0;     // first row of this launch (0: the whole array)
    int xcd;            // number workgroups per XCD (neighbouring rows share an L2)
    int slab;           // xcd with slabs: workgroups per plane in one XCD's slab (0: none)
    long long nrows;    // one past its last row (ny*nz: the whole array)
    int nx, ny, rows_per_wg, pitch;
    double norm;
    double *norm_part;  // NORM kernels: [2 * workgroups] (max |value|, max |result|)
    // refined form, lowest modes in FP64 (null: off).  x forward: low_rx[2 row
    // + kx] = sum_i x_i exp(-2 pi i kx i/nx), kx = 0, 1, of the FP64 row; x
    // backward: low_x[6 k + q] the solved modes of plane k, q = 3 kx + (0, 1,
    // 2 for ky = 0, 1, ny - 1), added to the row (normalised)
    double2 *low_rx;
    const double2 *low_x;
    const double2 *twx64, *twy64;  // exp(-2 pi i t/nx), exp(-2 pi i t/ny) in FP64
};

// Workgroups are dispatched to the 8 XCDs round-robin; renumbering them so
// that each XCD runs one contiguous range of tiles puts neighbouring column
// tiles (which share 128-byte lines when a tile row is narrower than a
// line) on one L2.
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned b = blockIdx.x, n = gridDim.x;
    return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
}

// First row of an x kernel's workgroup (row = k ny + j).  With a.slab, XCD x
// (= blockIdx & 7) takes the rows j in [x ny/8, (x+1) ny/8) of every plane,
// plane after plane: a stencil source reads rows j +- 1 and planes k +- 1 of
// each row, and those were read by the same XCD a slab (ny/8 rows) earlier,
// so they are still in its L2.  A contiguous range of rows per XCD puts a
// whole plane (about 6 MB of traffic at ny = 384) between a row and its z
// neighbours, more than the 4 MB L2; round-robin splits the y neighbours
// over XCDs.  Needs row0 = 0, nrows = ny nz and ny % (8 rows_per_wg) = 0.
template <typename V>
__device__ __forceinline__ long long x_row_first(const XArgs<V> &a)
{
    if (a.slab) {
        const unsigned b = blockIdx.x, x = b & 7, p = b >> 3;
        const unsigned k = p / (unsigned)a.slab, q = p - k * (unsigned)a.slab;
        return (long long)k * a.ny + (long long)(x * (unsigned)a.slab + q) * a.rows_per_wg;
    }
    return a.row0 + (long long)(a.xcd ? xcd_block() : blockIdx.x) * a.rows_per_wg;
}

// Sources that read neighbouring rows and planes (stencils) take the slab
// order (Src::kStencil); the others keep contiguous ranges.
template <class S, class = void>
struct is_stencil {
    static constexpr bool value = false;
};
template <class S>
struct is_stencil<S, decltype((void)S::kStencil)> {
    static constexpr bool value = S::kStencil;
};

template <typename V>
struct YArgs {
    pll_fftplan P;
    const V *tw;        // [ny] exp(-2 pi i t/ny)
    int L, ny, cols, ntiles, pitch;
};

// FP64 kernels with 12 or more points per thread need more than 128 VGPRs,
// so their workgroups are capped at 256 threads; FP32 values take half the
// registers.
template <int E, typename V = double2>
constexpr int max_threads()
{
    return (E >= 12 && sizeof(V) == sizeof(double2)) ? 256 : 512;
}
template <class F, typename V = double2>
constexpr int max_threads_f()
{
    return max_threads<F::E, V>();
}

// x forward: real rows -> packed half spectrum rows of W.
// Recombination twiddle w^m, m = t + e T: w^t (one load per thread) times
// w^(eT) (the same for the whole workgroup, so read through the scalar unit).
template <typename V>
__device__ __forceinline__ V tw_m(const V *__restrict__ tw, V wt, int e, int T)
{
    return e == 0 ? wt : cmul(wt, tw[e * T]);
}

template <class F, typename V, class Src>
__global__ void __launch_bounds__((max_threads_f<F, V>()))
k_xfwd(Src src, V *__restrict__ W, const V *__restrict__ twr, XArgs<V> a)
{
    using S = typename Sc<V>::T;
    extern __shared__ unsigned char lds_raw[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), L = F::length(a.P);
    const int rl = threadIdx.x / T, t = threadIdx.x - rl * T;
    const long long row = x_row_first(a) + rl;
    const bool valid = row < a.nrows;
    unsigned char *ml = lds_raw + (size_t)rl * a.pitch * lds_elem<F, V>();
    const long long j = valid ? row % a.ny : 0, k = valid ? row / a.ny : 0;
    V v[E];
    if constexpr (sizeof(S) == sizeof(float) && Src::kLow) {
        if (a.low_rx) {
            // FP64 sums of the row for kx = 0 and 1: each FP64 value x is
            // split into its FP32 rounding v (the transform's input) and the
            // rounding error x - v; the errors (at most 2^-24 |x|) are summed
            // in FP32 as they are loaded, the v afterwards in FP64, so that the
            // FP64 values of the row segment need not stay in registers.
            // exp(-2 pi i x/nx) at x = 2m, 2m + 1 by recurrences (m steps by T).
            float e0 = 0.0f, e1r = 0.0f, e1i = 0.0f;
            {
                const double2 a0 = a.twx64[2 * t], a1 = a.twx64[2 * t + 1];
                const double2 as = a.twx64[(2 * T) % a.nx];
                float2 u0 = make_float2((float)a0.x, (float)a0.y);
                float2 u1 = make_float2((float)a1.x, (float)a1.y);
                const float2 us = make_float2((float)as.x, (float)as.y);
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int m = t + e * T;
                    double2 d = make_double2(0.0, 0.0);
                    if (valid) d = src.load_d(row, j, k, m, a.nx);
                    v[e] = mk<V>((S)d.x, (S)d.y);
                    const float r0 = (float)(d.x - (double)v[e].x);
                    const float r1 = (float)(d.y - (double)v[e].y);
                    e0 += r0 + r1;
                    e1r += r0 * u0.x + r1 * u1.x;
                    e1i += r0 * u0.y + r1 * u1.y;
                    u0 = make_float2(u0.x * us.x - u0.y * us.y, u0.x * us.y + u0.y * us.x);
                    u1 = make_float2(u1.x * us.x - u1.y * us.y, u1.x * us.y + u1.y * us.x);
                }
            }
            double2 w0 = a.twx64[2 * t], w1 = a.twx64[2 * t + 1];
            const double2 st = a.twx64[(2 * T) % a.nx];
            double s0 = (double)e0, s1r = (double)e1r, s1i = (double)e1i;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const double x0 = (double)v[e].x, x1 = (double)v[e].y;
                s0 += x0 + x1;
                s1r += x0 * w0.x + x1 * w1.x;
                s1i += x0 * w0.y + x1 * w1.y;
                w0 = make_double2(w0.x * st.x - w0.y * st.y, w0.x * st.y + w0.y * st.x);
                w1 = make_double2(w1.x * st.x - w1.y * st.y, w1.x * st.y + w1.y * st.x);
            }
            for (int o = T >> 1; o > 0; o >>= 1) {
                s0 += __shfl_xor(s0, o);
                s1r += __shfl_xor(s1r, o);
                s1i += __shfl_xor(s1i, o);
            }
            if (valid && t == 0) {
                a.low_rx[2 * row] = make_double2(s0, 0.0);
                a.low_rx[2 * row + 1] = make_double2(s1r, s1i);
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int m = t + e * T;
                v[e] = valid ? src.template load<V>(row, j, k, m, a.nx) : mk<V>(S(0), S(0));
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            v[e] = valid ? src.template load<V>(row, j, k, m, a.nx) : mk<V>(S(0), S(0));
        }
    }
    F::template run<-1, 2>(v, t, a.P, a.tw, ml);
    V zm[E];
    if constexpr (shfl_mirror<F>()) {
        mirror<F, V, E, true>(v, t, T, L, ml, zm);
    } else {
        __syncthreads();
        mirror<F, V, E>(v, t, T, L, ml, zm);
    }
    V *dst = W + row * L;
    const V wt = twr[t];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = t + e * T;
        const V Z = v[e];
        V X;
        if (m == 0) {
            X = mk<V>(Z.x + Z.y, Z.x - Z.y);  // (X_0, X_L)
        } else {
            const V Zc = cconj(zm[e]);
            const V Ev = cscale(cadd(Z, Zc), S(0.5));
            const V D = csub(Z, Zc);
            const V O = mk<V>(S(0.5) * D.y, S(-0.5) * D.x);  // -i D / 2
            X = cadd(Ev, cmul(tw_m(twr, wt, e, T), O));
        }
        if (valid) dst[m] = X;
    }
}

// x backward: packed half spectrum rows -> real rows * norm, stored by Dst;
// with NORM, also the workgroup's max |value| and max |result|.
template <class F, typename V, class Dst, bool NORM = false>
__global__ void __launch_bounds__((max_threads_f<F, V>()))
k_xbwd(const V *__restrict__ W, Dst dst, const V *__restrict__ twr, XArgs<V> a)
{
    using S = typename Sc<V>::T;
    extern __shared__ unsigned char lds_raw[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), L = F::length(a.P);
    const int rl = threadIdx.x / T, t = threadIdx.x - rl * T;
    const long long row = x_row_first(a) + rl;
    const bool valid = row < a.nrows;
    unsigned char *ml = lds_raw + (size_t)rl * a.pitch * lds_elem<F, V>();
    const long long j = valid ? row % a.ny : 0, k = valid ? row / a.ny : 0;
    const V *src = W + row * L;
    V v[E];
    // all row loads in flight before the first LDS write (a load feeding an
    // LDS write directly is otherwise waited for one at a time)
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? src[t + e * T] : mk<V>(S(0), S(0));
    V xm[E];
    mirror<F, V, E, shfl_mirror<F>()>(v, t, T, L, ml, xm);
    const V wt = twr[t];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int m = t + e * T;
        const V X = v[e];
        if (m == 0) {
            v[e] = mk<V>(X.x + X.y, X.x - X.y);
        } else {
            const V Xc = cconj(xm[e]);
            const V Sm = cadd(X, Xc);
            const V D = csub(X, Xc);
            const V Tt = cmul(cconj(tw_m(twr, wt, e, T)), D);
            v[e] = mk<V>(Sm.x - Tt.y, Sm.y + Tt.x);  // S + i Tt
        }
    }
    if constexpr (!shfl_mirror<F>()) __syncthreads();  // the mirror's LDS reads are done
    F::template run<+1, 2>(v, t, a.P, a.tw, ml);
    if constexpr (!NORM && Dst::kAligned && sizeof(S) == sizeof(double)) {
        // x_{2m+1} of each pair through LDS to the next lane: stores of
        // (x_{2m-1}, x_{2m}) at the even haloed slot 2m are 16-byte aligned;
        // x_0 and x_{nx-1} go alone (FP64 rows: the row's LDS holds L doubles)
        double *ls = reinterpret_cast<double *>(ml);
        __syncthreads();  // the FFT's last LDS reads are done
#pragma unroll
        for (int e = 0; e < E; ++e) ls[t + e * T] = (double)v[e].y * a.norm;
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int m = t + e * T;
                const double x0 = (double)v[e].x * a.norm;
                if (m == 0)
                    dst.row_one(j, k, 1, x0);
                else
                    dst.row_pair(j, k, 2 * m, ls[m - 1], x0);
                if (m == L - 1) dst.row_one(j, k, 2 * L, ls[m]);
            }
        }
    } else {
    // refined form, lowest modes solved in FP64: their contribution to row
    // (j, k) is norm (B + 2 Re(A exp(2 pi i x/nx))), with
    // B = X(0,0) + 2 Re(X(0,1) e), A = X(1,0) + X(1,1) e + X(1,-1) conj(e),
    // e = exp(2 pi i j/ny); the FP32 pipeline carries none of these modes
    double lb = 0.0, lar = 0.0, lai = 0.0;
    const bool low = sizeof(S) == sizeof(float) && a.low_x != nullptr;
    if (low && valid) {
        const double2 *X = a.low_x + 6 * k;
        const double2 ty = a.twy64[j];
        const double ec = ty.x, es = -ty.y;  // e = exp(+2 pi i j/ny)
        const double2 x00 = X[0], x01 = X[1], x10 = X[3], x11 = X[4], x1m = X[5];
        lb = (x00.x + 2.0 * (x01.x * ec - x01.y * es)) * a.norm;
        lar = (x10.x + (x11.x * ec - x11.y * es) + (x1m.x * ec + x1m.y * es)) * a.norm;
        lai = (x10.y + (x11.x * es + x11.y * ec) + (x1m.y * ec - x1m.x * es)) * a.norm;
    }
    // exp(-2 pi i x/nx) (cos, -sin) at x = 2m, 2m + 1, advanced by a
    // recurrence as m steps by T
    double2 tw0 = make_double2(1.0, 0.0), tw1 = tw0, tst = tw0;
    if (low) {
        tw0 = a.twx64[2 * t];
        tw1 = a.twx64[2 * t + 1];
        tst = a.twx64[(2 * T) % a.nx];
    }
    double lc0 = 0.0, lc1 = 0.0;  // contributions at 2m, 2m + 1 of the current e
    auto lowstep = [&]() {
        if (!low) return;
        lc0 = lb + 2.0 * (lar * tw0.x + lai * tw0.y);
        lc1 = lb + 2.0 * (lar * tw1.x + lai * tw1.y);
        tw0 = make_double2(tw0.x * tst.x - tw0.y * tst.y, tw0.x * tst.y + tw0.y * tst.x);
        tw1 = make_double2(tw1.x * tst.x - tw1.y * tst.y, tw1.x * tst.y + tw1.y * tst.x);
    };
    if constexpr (!NORM) {
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int m = t + e * T;
                lowstep();
                dst.store(row, j, k, m, a.nx, (double)v[e].x * a.norm + lc0,
                          (double)v[e].y * a.norm + lc1);
            }
        }
    } else {
        double nc = 0.0, np = 0.0;
        if (valid) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int m = t + e * T;
                lowstep();
                const double x0 = (double)v[e].x * a.norm + lc0;
                const double x1 = (double)v[e].y * a.norm + lc1;
                np = nmax(np, dst.store_n(row, j, k, m, a.nx, x0, x1));
                nc = nmax(nc, nmax(fabs(x0), fabs(x1)));
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            nc = nmax(nc, __shfl_xor(nc, o));
            np = nmax(np, __shfl_xor(np, o));
        }
        __syncthreads();  // the FFT's last LDS reads are done
        double *red = reinterpret_cast<double *>(lds_raw);
        const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
        if ((threadIdx.x & 63) == 0) {
            red[2 * wv] = nc;
            red[2 * wv + 1] = np;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 1; q < nw; ++q) {
                nc = nmax(nc, red[2 * q]);
                np = nmax(np, red[2 * q + 1]);
            }
            a.norm_part[2 * blockIdx.x] = nc;
            a.norm_part[2 * blockIdx.x + 1] = np;
        }
    }
    }
}

// out[0..1] = max over the n partial pairs (fixed order)
__global__ void __launch_bounds__(256) k_norm_final(const double *__restrict__ part, long long n,
                                                    double *__restrict__ out)
{
    __shared__ double red[512];
    const int t = threadIdx.x;
    double nc = 0.0, np = 0.0;
    for (long long q = t; q < n; q += 256) {
        nc = nmax(nc, part[2 * q]);
        np = nmax(np, part[2 * q + 1]);
    }
    red[2 * t] = nc;
    red[2 * t + 1] = np;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            red[2 * t] = nmax(red[2 * t], red[2 * (t + h)]);
            red[2 * t + 1] = nmax(red[2 * t + 1], red[2 * (t + h) + 1]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = red[0];
        out[1] = red[1];
    }
}

// y forward, in place on W: tile of `cols` adjacent columns of one k-plane.
template <class F, typename V>
__global__ void __launch_bounds__((max_threads_f<F, V>()))
k_yfwd(V *__restrict__ W, YArgs<V> a)
{
    using S = typename Sc<V>::T;
    extern __shared__ unsigned char lds_raw[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), ny = F::length(a.P), L = a.L;
    const unsigned b = xcd_block();
    const int tile = b % a.ntiles;
    const long long k = b / a.ntiles;
    const int c = threadIdx.x % a.cols, t = threadIdx.x / a.cols;
    const int mx = tile * a.cols + c;
    const bool valid = mx < L;
    V *col = W + k * (long long)ny * L + (valid ? mx : 0);
    unsigned char *ml = lds_raw + (size_t)c * a.pitch * lds_elem<F, V>();
    V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int my = t + e * T;
        v[e] = valid ? col[(long long)my * L] : mk<V>(S(0), S(0));
    }
    F::template run<-1, 1>(v, t, a.P, a.tw, ml);
    if (tile == 0) {
        // column 0 holds DFT_y(X_0 + i X_L): split into A = DFT_y(X_0) and
        // B = DFT_y(X_L), keeping one half of each Hermitian sequence.
        __syncthreads();
        V vm[E];
        mirror<F, V, E>(v, t, T, ny, ml, vm);
        if (c == 0) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int my = t + e * T;
                if (my == 0 || 2 * my == ny) continue;  // (A, B) = (Re C, Im C)
                if (2 * my < ny) {
                    v[e] = cscale(cadd(v[e], cconj(vm[e])), S(0.5));  // A_my
                } else {
                    const V D = csub(vm[e], cconj(v[e]));
                    v[e] = mk<V>(S(0.5) * D.y, S(-0.5) * D.x);  // B_{ny-my}
                }
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; ++e) col[(long long)(t + e * T) * L] = v[e];
    }
}

template <class F, typename V>
__global__ void __launch_bounds__((max_threads_f<F, V>()))
k_ybwd(V *__restrict__ W, YArgs<V> a)
{
    using S = typename Sc<V>::T;
    extern __shared__ unsigned char lds_raw[];
    constexpr int E = F::E;
    const int T = F::threads(a.P), ny = F::length(a.P), L = a.L;
    const unsigned b = xcd_block();
    const int tile = b % a.ntiles;
    const long long k = b / a.ntiles;
    const int c = threadIdx.x % a.cols, t = threadIdx.x / a.cols;
    const int mx = tile * a.cols + c;
    const bool valid = mx < L;
    V *col = W + k * (long long)ny * L + (valid ? mx : 0);
    unsigned char *ml = lds_raw + (size_t)c * a.pitch * lds_elem<F, V>();
    V v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int my = t + e * T;
        v[e] = valid ? col[(long long)my * L] : mk<V>(S(0), S(0));
    }
    if (tile == 0) {
        // rebuild C = A + i B over all my from the packed halves
        V vm[E];
        mirror<F, V, E>(v, t, T, ny, ml, vm);
        if (c == 0) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int my = t + e * T;
                if (my == 0 || 2 * my == ny) continue;
                if (2 * my < ny) {
                    const V B = vm[e];
                    v[e] = mk<V>(v[e].x - B.y, v[e].y + B.x);
                } else {
                    const V A = cconj(vm[e]);
                    const V B = cconj(v[e]);
                    v[e] = mk<V>(A.x - B.y, A.y + B.x);
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

template <typename V>
XArgs<V> x_args(const pll_ppe_dev *d, const void *tw)
{
    XArgs<V> a;
    a.P = d->px;
    a.tw = (const V *)tw;
    a.row0 = 0;
    a.xcd = 0;
    a.slab = 0;
    a.low_rx = nullptr;
    a.low_x = nullptr;
    a.twx64 = (const double2 *)d->tw_x;
    a.twy64 = (const double2 *)d->tw_y;
    a.nrows = (long long)d->ny * d->nz;
    a.nx = d->nx;
    a.ny = d->ny;
    a.rows_per_wg = d->x_rows_per_wg;
    a.pitch = d->px.L + 1;
    a.norm = d->normfactor;
    return a;
}

// Rows [row0, row1) of the ny*nz rows (row = k ny + j); row1 < 0: all rows.
template <class F, typename V, class Src>
int launch_xf(const pll_ppe_dev *d, const void *tw, Src src, void *W, hipStream_t s,
              long long row0 = 0, long long row1 = -1)
{
    XArgs<V> a = x_args<V>(d, tw);
    if (row1 >= 0) {
        if (row0 < 0 || row0 >= row1 || row1 > a.nrows) return -1;
        a.row0 = row0;
        a.nrows = row1;
    }
    const int threads = a.rows_per_wg * a.P.T;
    const long long blocks = (a.nrows - a.row0 + a.rows_per_wg - 1) / a.rows_per_wg;
    const size_t shm = (size_t)a.rows_per_wg * a.pitch * lds_elem<F, V>();
    if (threads > max_threads_f<F, V>() || shm > 65536 || blocks > 0x7fffffffLL) return -1;
    a.xcd = d->x_xcd >= 2 || (d->x_xcd == 1 && sizeof(V) == sizeof(float2));
    if (a.xcd && d->x_slab && (is_stencil<Src>::value || d->x_slab >= 2) && a.row0 == 0 &&
        a.nrows == (long long)d->ny * d->nz && d->ny % (8 * a.rows_per_wg) == 0)
        a.slab = d->ny / (8 * a.rows_per_wg);
    if (sizeof(V) == sizeof(float2) && d->low_on) {
        if (!Src::kLow) return -1;
        a.low_rx = (double2 *)d->low_rx;
    }
    hipLaunchKernelGGL((k_xfwd<F, V, Src>), dim3((unsigned)blocks), dim3(threads), shm, s, src,
                       (V *)W, a.tw, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <class F, typename V, class Dst, bool NORM = false>
int launch_xb(const pll_ppe_dev *d, const void *tw, const void *W, Dst dst, hipStream_t s,
              double *norm = nullptr, long long row0 = 0, long long row1 = -1)
{
    XArgs<V> a = x_args<V>(d, tw);
    if (row1 >= 0) {
        if (NORM || row0 < 0 || row0 >= row1 || row1 > a.nrows) return -1;
        a.row0 = row0;
        a.nrows = row1;
    }
    const int threads = a.rows_per_wg * a.P.T;
    const long long blocks = (a.nrows - a.row0 + a.rows_per_wg - 1) / a.rows_per_wg;
    const size_t shm = (size_t)a.rows_per_wg * a.pitch * lds_elem<F, V>();
    if (threads > max_threads_f<F, V>() || shm > 65536 || blocks > 0x7fffffffLL) return -1;
    a.xcd = d->x_xcd >= 2 || (d->x_xcd == 1 && sizeof(V) == sizeof(float2));
    if (sizeof(V) == sizeof(float2) && d->low_on) a.low_x = (const double2 *)d->low_x;
    if (NORM) {
        // norm[0..1]: result; norm[2 ..): partials, room for ny*nz workgroups
        if (!norm || blocks > a.nrows || shm < sizeof(double) * 2 * ((threads + 63) / 64))
            return -1;
        a.norm_part = norm + 2;
    }
    hipLaunchKernelGGL((k_xbwd<F, V, Dst, NORM>), dim3((unsigned)blocks), dim3(threads), shm, s,
                       (const V *)W, dst, a.tw, a);
    if (hipGetLastError() != hipSuccess) return -3;
    if (NORM) {
        hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, s, (const double *)(norm + 2),
                           blocks, norm);
        if (hipGetLastError() != hipSuccess) return -3;
    }
    return 0;
}

template <class F, typename V>
int launch_y(bool fwd, const pll_ppe_dev *d, const void *tw, void *W, hipStream_t s)
{
    YArgs<V> a;
    a.P = d->py;
    a.tw = (const V *)tw;
    a.L = d->L;
    a.ny = d->ny;
    a.cols = sizeof(V) == sizeof(double2) ? d->y_cols_per_wg : d->y_cols32_per_wg;
    a.ntiles = (d->L + a.cols - 1) / a.cols;
    a.pitch = d->ny + 1;
    const int threads = a.cols * a.P.T;
    const long long blocks = (long long)a.ntiles * d->nz;
    const size_t shm = (size_t)a.cols * a.pitch * lds_elem<F, V>();
    if (threads > max_threads_f<F, V>() || shm > 163840 || blocks > 0x7fffffffLL) return -1;
    // long columns (ny = 1024 in FP64: 16 KB each) need more than the default
    // 64 KB of LDS for a tile of 4 columns
    const void *fn = fwd ? (const void *)k_yfwd<F, V> : (const void *)k_ybwd<F, V>;
    if (shm > 65536 &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) != hipSuccess)
        return -3;
    if (fwd)
        hipLaunchKernelGGL((k_yfwd<F, V>), dim3((unsigned)blocks), dim3(threads), shm, s, (V *)W, a);
    else
        hipLaunchKernelGGL((k_ybwd<F, V>), dim3((unsigned)blocks), dim3(threads), shm, s, (V *)W, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// Compile-time plans for the transform lengths of the production grids
// (x: L = nx/2, y: L = ny).  For a length listed twice the first entry is the
// default (fewer registers, more waves per SIMD: faster on gfx950 as measured)
// and PLL_FFT_E selects the other.
#define PLL_CT_PLANS(X)                 \
    X(32, 8, 8, 4)                      \
    X(48, 24, 8, 6)                     \
    X(64, 8, 8, 8)                      \
    X(96, 24, 8, 4, 3)                  \
    X(128, 8, 8, 8, 2)                  \
    X(192, 12, 4, 4, 4, 3)              \
    X(192, 24, 8, 8, 3)                 \
    X(256, 8, 8, 8, 4)                  \
    X(384, 12, 4, 4, 4, 6)              \
    X(384, 24, 8, 8, 6)                 \
    X(512, 8, 8, 8, 8)                  \
    X(768, 12, 4, 4, 4, 4, 3)           \
    X(768, 24, 8, 8, 4, 3)              \
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
        case 5: return g(RtFFT<5>{});
        case 6: return g(RtFFT<6>{});
        case 7: return g(RtFFT<7>{});
        case 8: return g(RtFFT<8>{});
        case 12: return g(RtFFT<12>{});
        case 24: return g(RtFFT<24>{});
        default: return -5;
    }
}

}  // namespace

extern "C" int pll_fft_supported_E(int E)
{
    return E == 2 || E == 3 || E == 4 || E == 5 || E == 6 || E == 7 || E == 8 || E == 12 ||
           E == 24;
}

extern "C" int pll_fft_max_threads(int E) { return E >= 12 ? 256 : 512; }
extern "C" int pll_fft_max_threads32(int) { return 512; }

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
    if (pick < 0) return 0;
    p->L = L;
    p->E = plans[pick][1];
    p->T = L / p->E;
    int n = 0, prod = 1;
    for (int j = 2; j < 8 && plans[pick][j] > 0; j++) {
        p->rad[n++] = plans[pick][j];
        prod *= plans[pick][j];
    }
    p->nst = n;
    return prod == L;
}

// ---- FP64 solve
extern "C" int pll_launch_xfwd(const pll_ppe_dev *d, const double *p, void *W, hipStream_t s)
{
    const SrcHalo64 src{p, (long long)(d->nx + 2) * (d->ny + 2), d->nx + 2};
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xf<decltype(f), double2>(d, d->tw_x, src, W, s);
    });
}

// The chunked transforms of host-pointer solves are their own instantiations
// (same code), so that kernel traces keep them apart from whole-array ones.
struct SrcHalo64Rows : SrcHalo64 {};
struct DstHalo64Rows : DstHalo64 {};

extern "C" int pll_launch_xfwd_rows(const pll_ppe_dev *d, const double *p, void *W,
                                    long long row0, long long row1, hipStream_t s)
{
    SrcHalo64Rows src;
    src.p = p;
    src.off1 = (long long)(d->nx + 2) * (d->ny + 2);
    src.off2 = d->nx + 2;
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xf<decltype(f), double2>(d, d->tw_x, src, W, s, row0, row1);
    });
}

extern "C" int pll_launch_xfwd_div(const pll_ppe_dev *d, const pll_div_src *q, void *W,
                                   hipStream_t s)
{
    if (q->unfilled) {
        const SrcDivUnfilled src{*q, (long long)(d->nx + 2) * (d->ny + 2), d->nx + 2, d->nx, d->ny,
                                 d->nz};
        return dispatch_plan(d->px, [&](auto f) {
            return launch_xf<decltype(f), double2>(d, d->tw_x, src, W, s);
        });
    }
    const SrcDiv src{*q, (long long)(d->nx + 2) * (d->ny + 2), d->nx + 2};
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xf<decltype(f), double2>(d, d->tw_x, src, W, s);
    });
}

extern "C" int pll_launch_xbwd(const pll_ppe_dev *d, const void *W, double *p, hipStream_t s)
{
    const DstHalo64 dst{p, (long long)(d->nx + 2) * (d->ny + 2), d->nx + 2};
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xb<decltype(f), double2>(d, d->tw_x, W, dst, s);
    });
}

extern "C" int pll_launch_xbwd_rows(const pll_ppe_dev *d, const void *W, double *p,
                                    long long row0, long long row1, hipStream_t s)
{
    DstHalo64Rows dst;
    dst.p = p;
    dst.off1 = (long long)(d->nx + 2) * (d->ny + 2);
    dst.off2 = d->nx + 2;
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xb<decltype(f), double2>(d, d->tw_x, W, dst, s, nullptr, row0, row1);
    });
}

extern "C" int pll_launch_xbwd_fill(const pll_ppe_dev *d, const void *W, double *p, hipStream_t s)
{
    const DstHaloFill dst{p, (long long)(d->nx + 2) * (d->ny + 2), d->nx + 2, d->nx, d->ny, d->nz};
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xb<decltype(f), double2>(d, d->tw_x, W, dst, s);
    });
}

extern "C" int pll_launch_yfwd(const pll_ppe_dev *d, void *W, hipStream_t s)
{
    return dispatch_plan(d->py, [&](auto f) { return launch_y<decltype(f), double2>(true, d, d->tw_y, W, s); });
}

extern "C" int pll_launch_ybwd(const pll_ppe_dev *d, void *W, hipStream_t s)
{
    return dispatch_plan(d->py, [&](auto f) { return launch_y<decltype(f), double2>(false, d, d->tw_y, W, s); });
}

// ---- FP32 solve of the refined form
extern "C" int pll_launch_xfwd32(const pll_ppe_dev *d, int src_kind, const void *src, void *W32,
                                 hipStream_t s)
{
    if (src_kind == PLL_SRC_HALO64) {
        const SrcHalo64 in{(const double *)src, (long long)(d->nx + 2) * (d->ny + 2), d->nx + 2};
        return dispatch_plan(d->px, [&](auto f) {
            return launch_xf<decltype(f), float2>(d, d->tw_x32, in, W32, s);
        });
    }
    const SrcDense32 in{(const float *)src};
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xf<decltype(f), float2>(d, d->tw_x32, in, W32, s);
    });
}

extern "C" int pll_launch_xbwd32(const pll_ppe_dev *d, int dst_kind, const void *W32,
                                 double *p_halo, double *P_dense, const double *eps,
                                 hipStream_t s)
{
    const long long off1 = (long long)(d->nx + 2) * (d->ny + 2), off2 = d->nx + 2;
    return dispatch_plan(d->px, [&](auto f) {
        using F = decltype(f);
        switch (dst_kind) {
            case PLL_DST_HALO_SET:
                return launch_xb<F, float2>(d, d->tw_x32, W32, DstHalo64{p_halo, off1, off2}, s);
            case PLL_DST_DENSE_SET:
                return launch_xb<F, float2>(d, d->tw_x32, W32, DstDenseSet{P_dense}, s);
            case PLL_DST_DENSE_ADD:
                return launch_xb<F, float2>(d, d->tw_x32, W32, DstDenseAdd{P_dense}, s);
            case PLL_DST_HALO_FROM_DENSE:
                return launch_xb<F, float2>(d, d->tw_x32, W32,
                                            DstHaloFromDense{p_halo, off1, off2, P_dense, eps}, s);
            default: return -1;
        }
    });
}

extern "C" int pll_launch_xfwd32_copy(const pll_ppe_dev *d, const double *f_halo, double *F,
                                      void *W32, hipStream_t s)
{
    const SrcHalo64Copy in{f_halo, (long long)(d->nx + 2) * (d->ny + 2), d->nx + 2, F};
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xf<decltype(f), float2>(d, d->tw_x32, in, W32, s);
    });
}

extern "C" int pll_launch_xfwd32_resid(const pll_ppe_dev *d, const double *f, int f_halo,
                                       const double *P, int p_halo, void *W32, hipStream_t s)
{
    const long long h1 = (long long)(d->nx + 2) * (d->ny + 2), h2 = d->nx + 2;
    const long long n1 = (long long)d->nx * d->ny, n2 = d->nx;
    SrcResid in;
    in.f = f;
    in.P = P;
    in.fb = f_halo ? h1 + h2 + 1 : 0;
    in.f1 = f_halo ? h1 : n1;
    in.f2 = f_halo ? h2 : n2;
    in.pb = p_halo ? h1 + h2 + 1 : 0;
    in.p1 = p_halo ? h1 : n1;
    in.p2 = p_halo ? h2 : n2;
    in.a = d->a;
    in.c = d->c;
    in.d = d->dz;
    in.idx2 = d->idx2;
    in.idy2 = d->idy2;
    in.ny = d->ny;
    in.nz = d->nz;
    return dispatch_plan(d->px, [&](auto fp) {
        return launch_xf<decltype(fp), float2>(d, d->tw_x32, in, W32, s);
    });
}

extern "C" int pll_launch_xbwd32_norm(const pll_ppe_dev *d, const void *W32, double *p_halo,
                                      const double *eps, double *norm, hipStream_t s)
{
    const long long off1 = (long long)(d->nx + 2) * (d->ny + 2), off2 = d->nx + 2;
    return dispatch_plan(d->px, [&](auto f) {
        return launch_xb<decltype(f), float2, DstHaloAddN, true>(
            d, d->tw_x32, W32, DstHaloAddN{p_halo, off1, off2, eps}, s, norm);
    });
}

extern "C" int pll_launch_yfwd32(const pll_ppe_dev *d, void *W32, hipStream_t s)
{
    return dispatch_plan(d->py, [&](auto f) { return launch_y<decltype(f), float2>(true, d, d->tw_y32, W32, s); });
}

extern "C" int pll_launch_ybwd32(const pll_ppe_dev *d, void *W32, hipStream_t s)
{
    return dispatch_plan(d->py, [&](auto f) { return launch_y<decltype(f), float2>(false, d, d->tw_y32, W32, s); });
}
