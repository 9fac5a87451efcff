// Batched tridiagonal solve along z of the pressure Poisson solve (gfx950).
//
// Replaces ssolve_trid, src/src_c/mod_poisson_solver.c:311-359, with the
// pivots precomputed as in init_poisson_solver :245-276 (_PRECALC_TRID_COEFF).
// The reference stores one pivot per real unknown (N doubles); here pivots
// depend only on the eigenvalue lambda_x(ix) + lambda_y(iy), so the table
// holds one column per distinct (ix, iy): piv[k][iy][ix], ix in [0, L],
// iy in [0, ny/2].
//
// One thread owns one complex slot (mx, my) of W (two real systems); a
// wavefront covers 64 consecutive mx, so every k-row access is a coalesced
// 1 KiB segment.  Forward elimination and back substitution follow the
// reference's arithmetic: m_k = a_k / b_{k-1}; x_k -= m_k x_{k-1};
// x_{n-1} /= b_{n-1}; x_k = (x_k - c_k x_{k+1}) / b_k.
//
// Gauge: with Neumann walls the (kx,ky) = (0,0) system is singular.  The
// reference divides by its roundoff-sized last pivot (or writes 0 when the
// pivot is exactly 0, as on a uniform grid).  Here the top value of that
// mode is always fixed at 0, which is the reference's own choice whenever
// its last pivot is exactly zero; solutions differ from the reference's only
// by an additive constant.
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "pll_internal.h"

namespace {

typedef double2 cd;

struct TArgs {
    const double *a, *c, *piv;
    int L, ny, nz, pin_gauge;
};

template <typename C>
__global__ void __launch_bounds__(256) k_trid(C *__restrict__ W, TArgs g)
{
    const long long nslot = (long long)g.L * g.ny;
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int L = g.L, ny = g.ny, nz = g.nz;
    const int mx = (int)(s % L), my = (int)(s / L);
    const int iy = my <= ny - my ? my : ny - my;
    int ixr = mx, ixi = mx;
    if (mx == 0) {
        if (my == 0 || 2 * my == ny) {
            ixr = 0;
            ixi = L;
        } else if (2 * my < ny) {
            ixr = ixi = 0;
        } else {
            ixr = ixi = L;
        }
    }
    const long long pst = (long long)(L + 1) * (ny / 2 + 1);
    const double *__restrict__ pr = g.piv + (long long)iy * (L + 1) + ixr;
    const double *__restrict__ pi = g.piv + (long long)iy * (L + 1) + ixi;
    const bool same = (ixr == ixi);
    const double *__restrict__ a = g.a;
    const double *__restrict__ c = g.c;
    C *__restrict__ col = W + s;

    cd prev = make_double2(col[0].x, col[0].y);
    for (int k = 1; k < nz; ++k) {
        cd cur = make_double2(col[(long long)k * nslot].x, col[(long long)k * nslot].y);
        const double mr = a[k] / pr[(long long)(k - 1) * pst];
        const double mi = same ? mr : a[k] / pi[(long long)(k - 1) * pst];
        cur.x -= mr * prev.x;
        cur.y -= mi * prev.y;
        col[(long long)k * nslot].x = cur.x;
        col[(long long)k * nslot].y = cur.y;
        prev = cur;
    }
    {
        const double br = pr[(long long)(nz - 1) * pst];
        const double bi = pi[(long long)(nz - 1) * pst];
        const bool pin = g.pin_gauge && s == 0;
        prev.x = pin ? 0.0 : (br != 0.0 ? prev.x / br : 0.0);
        prev.y = bi != 0.0 ? prev.y / bi : 0.0;
        col[(long long)(nz - 1) * nslot].x = prev.x;
        col[(long long)(nz - 1) * nslot].y = prev.y;
    }
    for (int k = nz - 2; k >= 0; --k) {
        cd cur = make_double2(col[(long long)k * nslot].x, col[(long long)k * nslot].y);
        const double br = pr[(long long)k * pst];
        const double bi = same ? br : pi[(long long)k * pst];
        cur.x = (cur.x - c[k] * prev.x) / br;
        cur.y = (cur.y - c[k] * prev.y) / bi;
        col[(long long)k * nslot].x = cur.x;
        col[(long long)k * nslot].y = cur.y;
        prev = cur;
    }
}


// ---------------------------------------------------------------------------
// Partitioned solve: one workgroup owns a tile of S consecutive real columns
// (q0 .. q0+S-1 of one my-row, 8*S contiguous bytes per k) over all nz rows.
// The tile is staged through LDS with coalesced row loads; then P = nz/M
// lanes per column each hold M consecutive rows in registers:
//   1. local elimination ("modified Thomas"): every row of the lane's block
//      is rewritten as  a'_i x_first + x_i + c'_i x_last = d'_i, and the
//      block's first/last rows couple only to the neighbouring blocks' last/
//      first unknowns;
//   2. the 2P interface unknowns (first, last of every block) form a
//      tridiagonal system with unit diagonal, solved by parallel cyclic
//      reduction across the P lanes with wavefront shuffles (the P lanes of
//      a column sit in one wavefront);
//   3. back substitution x_i = d'_i - a'_i x_first - c'_i x_last.
// Every value is read from and written to HBM once; no pivot table.
// Columns whose eigenvalue is close to the singular end of the spectrum are
// more accurate with the sequential recurrence (the block decomposition loses
// up to a few 1e-12 there), so tiles holding such columns run the reference's
// Thomas recurrence with the precomputed pivots on the LDS-resident column
// instead; the host lists them first in `order` so they start first.
struct PArgs {
    const double *a, *c, *bz, *lamx, *lamy, *piv;
    const int *order;
    int L, nx, ny, nz, P, lP, S, ntx, ntiles, pin_gauge;  // P = 1 << lP
    int qx, iyx;  // exact path for tiles with q0 < qx and iy < iyx
};

// 1/x from the hardware reciprocal plus two Newton steps (full double
// precision, not necessarily correctly rounded); x is a nonzero pivot.
__device__ __forceinline__ double rcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Lane exchange within 16-lane rows (DPP): ROW_SHR(n) gives lane l the value
// of lane l-n, ROW_SHL(n) that of lane l+n, and 0 when that lane lies outside
// l's row (the `old` operand is 0 and bound_ctrl is set).  With the 16 lanes
// of a column forming one row, a missing neighbour is read as the empty
// equation without any select.
constexpr int ROW_SHL(int n) { return 0x100 + n; }
constexpr int ROW_SHR(int n) { return 0x110 + n; }
template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

template <int T>
__device__ __forceinline__ void pcr16_stage(double &fa, double &fg, double &fd, double &la,
                                            double &lg, double &ld)
{
    const double pfa = dpp<ROW_SHR(T)>(fa), pfg = dpp<ROW_SHR(T)>(fg), pfd = dpp<ROW_SHR(T)>(fd);
    const double nfa = dpp<ROW_SHL(T)>(fa), nfg = dpp<ROW_SHL(T)>(fg), nfd = dpp<ROW_SHL(T)>(fd);
    const double pla = dpp<ROW_SHR(T)>(la), plg = dpp<ROW_SHR(T)>(lg), pld = dpp<ROW_SHR(T)>(ld);
    const double nla = dpp<ROW_SHL(T)>(la), nlg = dpp<ROW_SHL(T)>(lg), nld = dpp<ROW_SHL(T)>(ld);
    const double rf = rcp(1.0 - fa * pfg - fg * nfa);
    const double fa2 = -fa * pfa * rf, fg2 = -fg * nfg * rf;
    const double fd2 = (fd - fa * pfd - fg * nfd) * rf;
    const double rl = rcp(1.0 - la * plg - lg * nla);
    const double la2 = -la * pla * rl, lg2 = -lg * nlg * rl;
    const double ld2 = (ld - la * pld - lg * nld) * rl;
    fa = fa2; fg = fg2; fd = fd2; la = la2; lg = lg2; ld = ld2;
}

// LDS slot of row k, column c of a tile SW = S H columns wide: the column
// within its group of S is swizzled by the row's block k / M, so that the P
// lanes of a column (rows p M + i) fall in distinct banks.
__device__ __forceinline__ int swz(int k, int c, int S, int SW, int M)
{
    return k * SW + (c & ~(S - 1)) + ((c & (S - 1)) ^ ((k / M) & (S - 1)));
}

__device__ __forceinline__ double shfl(double v, int src)
{
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// eigenvalue column index of real column q (slot q/2, part q%2) in row my
__device__ __forceinline__ int lam_ix(int q, int my, int L, int ny)
{
    const int mx = q >> 1, part = q & 1;
    if (mx > 0) return mx;
    if (my == 0 || 2 * my == ny) return part ? L : 0;
    return (2 * my < ny) ? 0 : L;
}

// Persistent: workgroup b solves tiles order[b], order[b + G], ... (G the
// grid), loading the next tile's values into registers while it solves the
// current one, so the HBM reads of one tile overlap the arithmetic of the
// previous.  With G a multiple of 8, position t of the order still runs on
// XCD t % 8.
// DPP16: the P = 16 lanes of a column form one DPP row; the cyclic reduction
// then exchanges neighbours with DPP instead of ds_bpermute + select.
// H > 1 (FP32 storage): a tile is H groups of S columns, SW = S H columns
// wide in memory (256-byte row segments where S floats make only 128), and
// the S P threads solve the groups one after the other; g.ntx counts tiles of
// SW columns.
template <int M, int S, typename St, bool DPP16, int H = 1>
__global__ void __launch_bounds__(512) k_trid_part(St *__restrict__ Wr,
                                                   const int *__restrict__ order, PArgs g)
{
    extern __shared__ double tl[];
    const int lP = g.lP, P = 1 << lP, nz = g.nz;
    constexpr int SW = S * H;
    constexpr int LSW = SW == 64 ? 6 : (SW == 32 ? 5 : (SW == 16 ? 4 : (SW == 8 ? 3 : 2)));
    static_assert(S == 4 || S == 8 || S == 16 || S == 32, "columns per tile");
    static_assert(H == 1 || H == 2, "column groups per tile");
    const int tid = threadIdx.x;
    const long long plane = (long long)g.nx * g.ny;
    const int nthr = S * P;
    // z coefficients staged in LDS after the tile: row
    // k = p M + i of (a, c, bz) at cf[p CS + 3 i + {0, 1, 2}].  The odd pitch
    // CS = 3M + 1 puts the P lanes of a column in distinct banks, and a lane's
    // reads are at fixed offsets from one address.
    constexpr int CS = 3 * M + 1;
    double *cf = tl + (size_t)nz * SW;

    for (int k = tid; k < nz; k += nthr) {
        const int o = (k / M) * CS + 3 * (k % M);
        cf[o] = g.a[k];
        cf[o + 1] = g.c[k];
        cf[o + 2] = g.bz[k];
    }
    // element j of this thread: column st, row kt + j P/H (nthr = SW P/H),
    // at tile_base + toff + j jstride
    const int lPH = lP - (H == 2), st = tid & (SW - 1), kt = tid >> LSW;
    const long long toff = st + kt * plane, jstride = (long long)(P / H) * plane;
    St pre[H * M];
    auto fetch = [&](int tile) {
        const St *__restrict__ src =
            Wr + (long long)(tile / g.ntx) * g.nx + (tile % g.ntx) * SW + toff;
#pragma unroll
        for (int j = 0; j < H * M; ++j) pre[j] = src[j * jstride];
    };
    const int s = tid >> lP, p = tid & (P - 1);
    // eigenvalues of this thread's column in each group of a tile, loaded
    // with the tile's values (so one wait covers both)
    double lams[H];
    auto load_lams = [&](int tile) {
        const int my = tile / g.ntx, qt0 = (tile % g.ntx) * SW;
        const int iy = my <= g.ny - my ? my : g.ny - my;
#pragma unroll
        for (int h = 0; h < H; ++h)
            lams[h] = g.lamx[lam_ix(qt0 + h * S + s, my, g.L, g.ny)] + g.lamy[iy];
    };
    int tile = (int)blockIdx.x < g.ntiles ? order[blockIdx.x] : 0;
    if ((int)blockIdx.x < g.ntiles) {
        fetch(tile);
        load_lams(tile);
    }
    // Order of the global accesses in an iteration: wait for this tile's
    // values (and eigenvalues), issue the next tile's loads, solve, then store
    // this tile's results straight from LDS.  The next iteration's wait for
    // its values leaves those stores in flight (loads and stores share one
    // in-order counter: the stores were issued after the loads waited for).
    // The prefetch is unconditional (the last iteration re-reads its own
    // tile), so that no wait depends on whether a next tile exists.
    for (int t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
    const int tn = t + (int)gridDim.x;
    const int tile_next = order[tn < g.ntiles ? tn : t];
    const int my = tile / g.ntx, qt0 = (tile % g.ntx) * SW;
    St *__restrict__ base = Wr + (long long)my * g.nx + qt0;
    const int iy = my <= g.ny - my ? my : g.ny - my;
    __syncthreads();  // the previous tile's results have been read from LDS
#pragma unroll
    for (int j = 0; j < H * M; ++j) tl[swz(kt + (j << lPH), st, S, SW, M)] = (double)pre[j];
    __syncthreads();
    double lamc[H];
#pragma unroll
    for (int h = 0; h < H; ++h) lamc[h] = lams[h];
    fetch(tile_next);
    load_lams(tile_next);

#pragma unroll 1
    for (int h = 0; h < H; ++h) {
    const int q = qt0 + h * S + s;
    const double lam = (H == 1 || h == 0) ? lamc[0] : lamc[H - 1];
    if (qt0 < g.qx && iy < g.iyx) {
        // The reference's recurrences with the stored pivots (low modes, where
        // the block elimination below loses accuracy):
        //   forward  y_k = d_k - m_k y_{k-1},  m_k = a_k / piv_{k-1},
        //   backward x_k = (y_k - c_k x_{k+1}) / piv_k,  x_{nz-1} from the top rule.
        // Lane p of a column owns rows pM .. pM+M-1.  Each recurrence is affine
        // in the value entering the lane, so a scan of the lanes' affine maps
        // over the P lanes of the column gives every lane its entering value,
        // and the lane then runs the recurrence over its rows as written.
        const int k0 = p * M;
        const int pb = k0 * SW + h * S + (s ^ (p & (S - 1)));
        const int lane = tid & 63;
        const long long pst = (long long)(g.L + 1) * (g.ny / 2 + 1);
        const double *__restrict__ pc =
            g.piv + (long long)iy * (g.L + 1) + lam_ix(q, my, g.L, g.ny);
        const double *__restrict__ cl = cf + p * CS;
        // the lane's rows stay in its LDS slots (registers hold the pivots)
        double pv[M];
#pragma unroll
        for (int i = 0; i < M; ++i) pv[i] = pc[(long long)(k0 + i) * pst];
        const double pleft = shfl(pv[M - 1], p > 0 ? lane - 1 : lane);  // pivot of row k0-1
        double al = 0.0, be = 1.0;  // lane map: y(last row) = al + be * y(row k0-1)
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double m = (k0 + i > 0) ? cl[3 * i] / (i ? pv[i - 1] : pleft) : 0.0;
            al = tl[pb + i * SW] - m * al;
            be = -m * be;
        }
        for (int t = 1; t < P; t *= 2) {
            const bool h = p >= t;
            const double al2 = shfl(al, h ? lane - t : lane), be2 = shfl(be, h ? lane - t : lane);
            if (h) {
                al = al + be * al2;
                be = be * be2;
            }
        }
        double prev = shfl(al, p > 0 ? lane - 1 : lane);  // y(row k0-1)
        if (p == 0) prev = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double m = (k0 + i > 0) ? cl[3 * i] / (i ? pv[i - 1] : pleft) : 0.0;
            const double yi = tl[pb + i * SW] - m * prev;
            tl[pb + i * SW] = yi;
            prev = yi;
        }
        // top row: pinned gauge, y / piv, or 0 for an exactly zero pivot
        const bool top = p == P - 1;
        const bool pin = g.pin_gauge && my == 0 && q == 0;
        const double bl = pv[M - 1];
        const double xtop = pin ? 0.0 : (bl != 0.0 ? prev / bl : 0.0);
        double ga = 0.0, de = 1.0;  // lane map: x(row k0) = ga + de * x(row k0+M)
#pragma unroll
        for (int i = M - 1; i >= 0; --i) {
            if (top && i == M - 1) {
                ga = xtop;
                de = 0.0;
            } else {
                const double c = cl[3 * i + 1];
                ga = (tl[pb + i * SW] - c * ga) / pv[i];
                de = -(c * de) / pv[i];
            }
        }
        for (int t = 1; t < P; t *= 2) {
            const bool h = p + t < P;
            const double ga2 = shfl(ga, h ? lane + t : lane), de2 = shfl(de, h ? lane + t : lane);
            if (h) {
                ga = ga + de * ga2;
                de = de * de2;
            }
        }
        double next = shfl(ga, top ? lane : lane + 1);  // x(row k0+M)
#pragma unroll
        for (int i = M - 1; i >= 0; --i) {
            const double x =
                (top && i == M - 1) ? xtop : (tl[pb + i * SW] - cl[3 * i + 1] * next) / pv[i];
            tl[pb + i * SW] = x;
            next = x;
        }
    } else {
        const int k0 = p * M;
        // swz(k0 + i, h S + s) = pb + i SW for i < M, since (k0 + i) / M = p
        const int pb = k0 * SW + h * S + (s ^ (p & (S - 1)));
        double d[M], ap[M], cp[M];
#pragma unroll
        for (int i = 0; i < M; ++i) d[i] = tl[pb + i * SW];
        // 1. local elimination
        const double *__restrict__ cl = cf + p * CS;
#define CA(i) cl[3 * (i)]
#define CC(i) cl[3 * (i) + 1]
#define CB(i) cl[3 * (i) + 2]
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const double r = rcp(CB(i) + lam);
            d[i] *= r;
            ap[i] = CA(i) * r;
            cp[i] = CC(i) * r;
        }
#pragma unroll
        for (int i = 2; i < M; ++i) {
            const double ai = CA(i);
            const double r = rcp((CB(i) + lam) - ai * cp[i - 1]);
            d[i] = r * (d[i] - ai * d[i - 1]);
            ap[i] = -r * ai * ap[i - 1];
            cp[i] = r * CC(i);
        }
#undef CA
#undef CC
#undef CB
#pragma unroll
        for (int i = M - 3; i >= 1; --i) {
            d[i] -= cp[i] * d[i + 1];
            ap[i] -= cp[i] * ap[i + 1];
            cp[i] = -cp[i] * cp[i + 1];
        }
        {
            const double r = rcp(1.0 - cp[0] * ap[1]);
            d[0] = r * (d[0] - cp[0] * d[1]);
            ap[0] = r * ap[0];
            cp[0] = -r * cp[0] * cp[1];
        }
        // 2. reduced system: equation 2p = block's first row (F), 2p+1 = last (L)
        double fa = ap[0], fg = cp[0], fd = d[0];
        double la = ap[M - 1], lg = cp[M - 1], ld = d[M - 1];
        if (g.pin_gauge && my == 0 && q == 0 && p == P - 1) {
            la = 0.0; lg = 0.0; ld = 0.0;  // top value of the (0,0) mode = 0
            fg = 0.0;
        }
        if (DPP16) {
            // stride 1: F_p <- L_{p-1}, L_p;  L_p <- F_p, F_{p+1}
            const double pla = dpp<ROW_SHR(1)>(la), plg = dpp<ROW_SHR(1)>(lg), pld = dpp<ROW_SHR(1)>(ld);
            const double nfa = dpp<ROW_SHL(1)>(fa), nfg = dpp<ROW_SHL(1)>(fg), nfd = dpp<ROW_SHL(1)>(fd);
            const double rf = rcp(1.0 - fa * plg - fg * la);
            const double fa2 = -fa * pla * rf, fg2 = -fg * lg * rf;
            const double fd2 = (fd - fa * pld - fg * ld) * rf;
            const double rl = rcp(1.0 - la * fg - lg * nfa);
            const double la2 = -la * fa * rl, lg2 = -lg * nfg * rl;
            const double ld2 = (ld - la * fd - lg * nfd) * rl;
            fa = fa2; fg = fg2; fd = fd2; la = la2; lg = lg2; ld = ld2;
            pcr16_stage<1>(fa, fg, fd, la, lg, ld);
            pcr16_stage<2>(fa, fg, fd, la, lg, ld);
            pcr16_stage<4>(fa, fg, fd, la, lg, ld);
            pcr16_stage<8>(fa, fg, fd, la, lg, ld);
        } else {
        const int lane = tid & 63;
        // stride 1: F_p <- L_{p-1}, L_p;  L_p <- F_p, F_{p+1}
        {
            const bool hasl = p > 0, hasr = p < P - 1;
            const int sl = hasl ? lane - 1 : lane, sr = hasr ? lane + 1 : lane;
            // every lane takes part in every shuffle; out-of-range neighbours
            // are then replaced by the empty equation
            const double ula = shfl(la, sl), ulg = shfl(lg, sl), uld = shfl(ld, sl);
            const double dfa = shfl(fa, sr), dfg = shfl(fg, sr), dfd = shfl(fd, sr);
            const double pla = hasl ? ula : 0.0, plg = hasl ? ulg : 0.0, pld = hasl ? uld : 0.0;
            const double nfa = hasr ? dfa : 0.0, nfg = hasr ? dfg : 0.0, nfd = hasr ? dfd : 0.0;
            // F: left = L_{p-1}, right = L_p
            const double rf = rcp(1.0 - fa * plg - fg * la);
            const double fa2 = -fa * pla * rf, fg2 = -fg * lg * rf;
            const double fd2 = (fd - fa * pld - fg * ld) * rf;
            // L: left = F_p, right = F_{p+1}
            const double rl = rcp(1.0 - la * fg - lg * nfa);
            const double la2 = -la * fa * rl, lg2 = -lg * nfg * rl;
            const double ld2 = (ld - la * fd - lg * nfd) * rl;
            fa = fa2; fg = fg2; fd = fd2; la = la2; lg = lg2; ld = ld2;
        }
        // stride 2t: F_p <- F_{p-t}, F_{p+t};  L_p <- L_{p-t}, L_{p+t}
        for (int t = 1; t < P; t *= 2) {
            const bool hasl = p - t >= 0, hasr = p + t < P;
            const int sl = hasl ? lane - t : lane, sr = hasr ? lane + t : lane;
            const double ufa = shfl(fa, sl), ufg = shfl(fg, sl), ufd = shfl(fd, sl);
            const double dfa = shfl(fa, sr), dfg = shfl(fg, sr), dfd = shfl(fd, sr);
            const double ula = shfl(la, sl), ulg = shfl(lg, sl), uld = shfl(ld, sl);
            const double dla = shfl(la, sr), dlg = shfl(lg, sr), dld = shfl(ld, sr);
            const double pfa = hasl ? ufa : 0.0, pfg = hasl ? ufg : 0.0, pfd = hasl ? ufd : 0.0;
            const double nfa = hasr ? dfa : 0.0, nfg = hasr ? dfg : 0.0, nfd = hasr ? dfd : 0.0;
            const double pla = hasl ? ula : 0.0, plg = hasl ? ulg : 0.0, pld = hasl ? uld : 0.0;
            const double nla = hasr ? dla : 0.0, nlg = hasr ? dlg : 0.0, nld = hasr ? dld : 0.0;
            const double rf = rcp(1.0 - fa * pfg - fg * nfa);
            const double fa2 = -fa * pfa * rf, fg2 = -fg * nfg * rf;
            const double fd2 = (fd - fa * pfd - fg * nfd) * rf;
            const double rl = rcp(1.0 - la * plg - lg * nla);
            const double la2 = -la * pla * rl, lg2 = -lg * nlg * rl;
            const double ld2 = (ld - la * pld - lg * nld) * rl;
            fa = fa2; fg = fg2; fd = fd2; la = la2; lg = lg2; ld = ld2;
        }
        }
        // 3. back substitution
        const double F = fd, Lv = ld;
#pragma unroll
        for (int i = 1; i < M - 1; ++i) d[i] = d[i] - ap[i] * F - cp[i] * Lv;
        d[0] = F;
        d[M - 1] = Lv;
#pragma unroll
        for (int i = 0; i < M; ++i) tl[pb + i * SW] = d[i];
    }
    }  // h
    __syncthreads();
    St *__restrict__ dst = base + toff;
#pragma unroll
    for (int j = 0; j < H * M; ++j) dst[j * jstride] = (St)tl[swz(kt + (j << lPH), st, S, SW, M)];
    tile = tile_next;
    }
}

// Workgroups of a persistent launch: as many as are co-resident (PLL_TRID_WG
// per CU overrides), a multiple of 8, at most one per tile.
unsigned persistent_grid(const void *fn, int threads, size_t shm, int ntiles)
{
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            ncu <= 0)
            ncu = 256;
    }
    int per_cu = 0;
    const char *e = getenv("PLL_TRID_WG");
    if (e) per_cu = atoi(e);
    if (per_cu <= 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, shm) != hipSuccess ||
         per_cu <= 0))
        per_cu = 1;
    long long n = (long long)per_cu * ncu;
    n -= n % 8;
    if (n < 8) n = 8;
    return (unsigned)(n < ntiles ? n : ntiles);
}

template <typename St, typename C>
int launch_trid(const pll_ppe_dev *d, void *W, hipStream_t s)
{
    if (d->tp.P > 0) {
        // FP32 storage with wide tiles: H = 2 groups of S columns per tile
        const int H = ((sizeof(St) == sizeof(float) ? d->tp.h32 : d->tp.h64) == 2) ? 2 : 1;
        // H = 2: FP32 storage at S = 16 and 32, FP64 storage at S = 16 only (FP64
        // values of a 64-column tile need more registers than a wave has)
        constexpr int HW32 = sizeof(St) == sizeof(float) ? 2 : 1;
        PArgs g;
        g.a = d->a;
        g.c = d->c;
        g.bz = d->tp.bz;
        g.lamx = d->tp.lamx;
        g.lamy = d->tp.lamy;
        g.piv = d->piv;
        g.order = H == 2 ? d->tp.order32 : d->tp.order;
        g.L = d->L;
        g.nx = d->nx;
        g.ny = d->ny;
        g.nz = d->nz;
        g.P = d->tp.P;
        g.lP = 0;
        while ((1 << g.lP) < g.P) g.lP++;
        if ((1 << g.lP) != g.P) return -1;
        g.S = d->tp.S;
        if (d->nx % (g.S * H) || (H == 2 && (g.S < 16 || g.P < 2 || !g.order))) return -1;
        g.ntx = d->nx / (g.S * H);
        g.ntiles = g.ntx * d->ny;
        g.pin_gauge = d->pin_gauge;
        g.qx = H == 2 ? d->tp.qx32 : d->tp.qx;
        g.iyx = d->tp.iyx;
        const int threads = g.S * g.P;
        // values + coefficients (P rows of 3M + 1)
        const size_t shm = sizeof(double) * ((size_t)d->nz * g.S * H + 3 * (size_t)d->nz + g.P);
        if (threads > 512 || shm > 163840) return -1;
        switch (d->tp.M) {
#define PLL_TRID_LAUNCH(M_, S_, D_, H_)                                                          \
    do {                                                                                          \
        if (shm > 65536 &&                                                                        \
            hipFuncSetAttribute((const void *)k_trid_part<M_, S_, St, D_, H_>,                    \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) != hipSuccess) \
            return -3;                                                                            \
        const unsigned blocks = persistent_grid((const void *)k_trid_part<M_, S_, St, D_, H_>,   \
                                                threads, shm, g.ntiles);                          \
        hipLaunchKernelGGL((k_trid_part<M_, S_, St, D_, H_>), dim3(blocks), dim3(threads), shm, s, \
                           (St *)W, g.order, g);                                                  \
    } while (0)
#define PLL_TRID_S(M_, D_)                                                                      \
    switch (g.S) {                                                                              \
        case 4: PLL_TRID_LAUNCH(M_, 4, D_, 1); break;                                           \
        case 8: PLL_TRID_LAUNCH(M_, 8, D_, 1); break;                                           \
        case 16:                                                                                \
            if (H == 2)                                                                         \
                PLL_TRID_LAUNCH(M_, 16, D_, 2);                                                \
            else                                                                                \
                PLL_TRID_LAUNCH(M_, 16, D_, 1);                                                 \
            break;                                                                              \
        case 32:                                                                                \
            if (H == 2)                                                                         \
                PLL_TRID_LAUNCH(M_, 32, D_, HW32);                                                \
            else                                                                                \
                PLL_TRID_LAUNCH(M_, 32, D_, 1);                                                 \
            break;                                                                              \
        default: return -5;                                                                     \
    }
#define PLL_TRID_CASE(M_)                                                                       \
    case M_:                                                                                    \
        if (d->tp.dpp16) {                                                                      \
            PLL_TRID_S(M_, true)                                                                \
        } else {                                                                                \
            PLL_TRID_S(M_, false)                                                               \
        }                                                                                       \
        break;
            PLL_TRID_CASE(4)
            PLL_TRID_CASE(8)
            PLL_TRID_CASE(12)
            PLL_TRID_CASE(16)
#undef PLL_TRID_CASE
#undef PLL_TRID_S
#undef PLL_TRID_LAUNCH
            default: return -5;
        }
        return hipGetLastError() == hipSuccess ? 0 : -3;
    }
    TArgs g;
    g.a = d->a;
    g.c = d->c;
    g.piv = d->piv;
    g.L = d->L;
    g.ny = d->ny;
    g.nz = d->nz;
    g.pin_gauge = d->pin_gauge;
    const long long nslot = (long long)d->L * d->ny;
    const long long blocks = (nslot + 255) / 256;
    if (blocks > 0x7fffffffLL) return -1;
    hipLaunchKernelGGL((k_trid<C>), dim3((unsigned)blocks), dim3(256), 0, s, (C *)W, g);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace

extern "C" int pll_trid_part_supported(int M) { return M == 4 || M == 8 || M == 12 || M == 16; }

extern "C" int pll_launch_trid(const pll_ppe_dev *d, void *W, hipStream_t s)
{
    return launch_trid<double, double2>(d, W, s);
}

extern "C" int pll_launch_trid32(const pll_ppe_dev *d, void *W32, hipStream_t s)
{
    return launch_trid<float, float2>(d, W32, s);
}
