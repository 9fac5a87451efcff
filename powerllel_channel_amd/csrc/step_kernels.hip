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

#include <stdlib.h>

#include <type_traits>

#include "mom_rhs.h"
#include "pll_internal.h"

namespace {

using pll::G;
using pll::Nb;
using pll::make_g;
using pll::mom_rhs;


// RK stencils march over a chunk of k per thread: the values of plane k+1
// that become plane k (and then k-1) at the next step stay in registers, so
// each plane of the three fields is fetched once per chunk; x and y
// neighbours come through the caches.  Workgroups are numbered so that each
// of the 8 XCDs (which take workgroups round-robin) gets one contiguous run
// of tiles: y-adjacent tiles of one k-chunk then share that XCD's L2.
constexpr int RK_TX = 64, RK_TY = 4;
struct Tiles {
    int ntx, nty, nkc, kc;
    int xcd;  // 1: remap workgroup ids per XCD (tile count divisible by 8)
};

__device__ __forceinline__ void rk_tile(const Tiles &t, int &tx, int &ty, int &kc)
{
    const int nb = t.ntx * t.nty * t.nkc;
    int lb = blockIdx.x;
    if (t.xcd) lb = (lb & 7) * (nb >> 3) + (lb >> 3);
    tx = lb % t.ntx;
    lb /= t.ntx;
    ty = lb % t.nty;
    kc = lb / t.nty;
}

// SECOND = false: o := q + dt f(q)                  (RK1, o = u1, q = u)
// SECOND = true:  o := (q + dt f(q) + o) * 0.5      (RK2, o = u,  q = u1)
// MEAN (RK2 only): also the Mean_h partial sums of the new u (fmask bit 0)
// and v (bit 1), sum of o * dzflzi(k) over the block's cells, into
// mpart[blockIdx.x] and mpart[nblocks + blockIdx.x].
//
// 2.5-D blocking: a workgroup owns a 64 x 4 column tile and marches over a
// k chunk.  Planes k and k+1 of the three fields sit in LDS as 66 x 6 tiles
// (tile plus one-cell halo), loaded with coalesced row reads; plane k+2 is
// fetched into registers while plane k is computed and written over plane
// k's buffer afterwards.  The few plane k-1 values the stencil reads are
// carried in registers from the previous step.  Every stencil read is then
// an LDS read, and each field value is loaded from memory about 1.5 times
// (the halo) instead of once per neighbour that reads it.
// BODY: a body force b (haloed layout, interior read) is added to the
// right-hand side, f := f(q) + b, in both stages.
struct Body {
    const double *u, *v, *w;
};

__device__ __forceinline__ int wrap(int i, int n) { return i == 0 ? n : (i == n + 1 ? 1 : i); }

// CORR (RK1 of a multi-step run): the velocity q is the previous step's RK2
// output after its boundary fill, not yet projected.  Every value the
// stencil reads is formed on load as correctVel (with the forcing shift) and
// updateBoundVel would have left it, in the same roundings (k_correct,
// k_bc_vel), from q and the pressure p with its ghosts; the projected
// interior values are also written to (nu, nv, nw).
struct Corr {
    const double *p;
    double *nu, *nv, *nw;
    double dtdxi, dtdyi;
    int fmask;
    double vfu, vfv;
    const double *mean;
};

// The projecting RK1 (CORR) uses tiles of 64 x 8 columns: its pressure tiles
// add to the halo, so taller tiles pay (768x384x256: 4.60 -> 4.47 ms per step),
// while the other RK kernels are faster with 64 x 4 (4.74 vs 4.82 ms).
constexpr int RK_TY_CORR = 8;
constexpr int LTX = RK_TX + 2;
// CORR: pressure tiles one cell wider on the high x and y sides (p(i+1),
// p(j+1) of every velocity tile cell)
constexpr int PTX = LTX + 1;

// periodic index of a pressure tile cell: 0 -> n, n + 1 -> 1, n + 2 -> 2
__device__ __forceinline__ int wrap_p(int i, int n) { return i == 0 ? n : (i > n ? i - n : i); }

// FILL: q's boundary fill has not been made; the values it would leave
// (k_bc_vel) are read from the interior instead of the ghost cells.
// FULL: every tile lies inside the grid (nx % 64 == 0, ny % TY == 0).  Then no
// lane is masked: every lane loads its plane share (lanes past the halo tile
// load and store its last element again) and the last planes of a chunk are
// fetched again instead of skipped, so the loop has no divergent branch and
// each wait covers only the loads it needs: the prefetch is waited for when it
// is written to LDS (the plane's stores may stay in flight), and RK2's old
// values of plane k are loaded before the prefetch of plane k+2, so that
// waiting for them leaves the prefetch in flight.
// Occupancy hints: the FULL RK1 kernels at 4 waves per SIMD; the general RK2
// kernels without a body force at 5 (96 VGPRs, no spills: the forced RK2 of
// pll_channel_run had 100 and 4 waves; step 4.365 -> 4.350 ms at 768x384x256);
// the PROJO RK2 at 4 (at 5 it spills).
// PROJO (RK2 of a multi-step run): o is the previous step's RK2 output, not
// yet projected; its projected value at the cell (correctVel with the forcing
// shift, and w = 0 on the top interior row as updateBoundVel leaves it) is
// formed as it is read, from o and p, in the roundings of the projecting RK1,
// so that RK1 need not write the projected field.
template <bool SECOND, bool MEAN, bool BODY, bool CORR = false, int TY = RK_TY, bool FILL = false,
          bool FULL = false, bool PROJO = false>
__global__ void __launch_bounds__(RK_TX *TY, ((FULL && !SECOND) || PROJO ? 4 : (SECOND && !FULL && !BODY ? 5 : 1))) k_rk(G g, Tiles t, const double *__restrict__ qu,
                                                     const double *__restrict__ qv,
                                                     const double *__restrict__ qw,
                                                     double *__restrict__ ou,
                                                     double *__restrict__ ov,
                                                     double *__restrict__ ow,
                                                     double *__restrict__ mpart, int fmask,
                                                     Body fb, Corr cr = Corr{})
{
    constexpr int LTY = TY + 2, LPL = LTX * LTY, NT = RK_TX * TY;
    constexpr int NLD = (LPL + NT - 1) / NT;
    constexpr int PTY = LTY + 1, LPP = PTX * PTY, NLP = (LPP + NT - 1) / NT;
    __shared__ double tl[2][3][LPL];
    int tx, ty, kc;
    rk_tile(t, tx, ty, kc);
    const int tid = threadIdx.x;
    const int lx = tid % RK_TX, ly = tid / RK_TX;
    const int i0 = tx * RK_TX, j0 = ty * TY;  // haloed index of the tile's halo corner
    const int i = i0 + lx + 1, j = j0 + ly + 1;
    const bool valid = FULL || (i <= g.nx && j <= g.ny);
    const int k0 = kc * t.kc + 1, k1 = min(k0 + t.kc - 1, g.nz);
    const long long sy = g.sy, sz = g.sz;
    const double *__restrict__ Q[3] = {qu, qv, qw};
    // this thread's share of a plane tile: element q = tid + r * 256
    long long go[NLD];
    bool gk[NLD];
    int gii[NLD], gjj[NLD];
#pragma unroll
    for (int r = 0; r < NLD; ++r) {
        const int q = FULL ? min(tid + r * NT, LPL - 1) : tid + r * NT;
        const int row = q / LTX, col = q - row * LTX;
        const int gi = i0 + col, gj = j0 + row;
        gk[r] = FULL || (q < LPL && gi <= g.nx + 1 && gj <= g.ny + 1);
        go[r] = gi + sy * gj;
        gii[r] = gk[r] ? wrap(gi, g.nx) : 1;
        gjj[r] = gk[r] ? wrap(gj, g.ny) : 1;
    }
    // CORR: projected value of field f at interior (i, j, k), as k_correct
    const double shu = ((CORR || PROJO) && (cr.fmask & 1)) ? cr.vfu - cr.mean[0] : 0.0;
    const double shv = ((CORR || PROJO) && (cr.fmask & 2)) ? cr.vfv - cr.mean[1] : 0.0;
    auto proj = [&](int f, int ii, int jj, int kk) -> double {
        const long long e0 = ii + sy * jj + sz * kk;
        const double pc = cr.p[e0];
        if (f == 0) {
            double un = qu[e0] - (cr.p[e0 + 1] - pc) * cr.dtdxi;
            if (cr.fmask & 1) un = un + shu;
            return un;
        }
        if (f == 1) {
            double vn = qv[e0] - (cr.p[e0 + sy] - pc) * cr.dtdyi;
            if (cr.fmask & 2) vn = vn + shv;
            return vn;
        }
        return qw[e0] - (cr.p[e0 + sz] - pc) * g.dt * g.dzc_inv[kk];
    };
    // ... and as k_bc_vel then leaves it at haloed (gi, gj, kk) (ii, jj wrapped)
    auto filled = [&](int f, int ii, int jj, int kk) -> double {
        if (kk == 0 || kk == g.nz + 1) {
            if (f == 2) return 0.0;
            const double bc = f == 0 ? 0.0 - g.u_crf : 0.0;
            return 2.0 * bc - proj(f, ii, jj, kk == 0 ? 1 : g.nz);
        }
        if (f == 2 && kk == g.nz) return 0.0;
        return proj(f, ii, jj, kk);
    };
    double pre[3][NLD];
    // CORR: the plane's raw velocity (at the wrapped source cells) is loaded
    // ahead like pre; the pressure goes through two LDS tiles, p of the source
    // plane and of the plane above (w's projection), the latter loaded ahead
    // into pnext; the projection is formed when the plane is written to LDS
    __shared__ double tp[CORR ? 2 : 1][CORR ? LPP : 1];
    long long gpo[NLP];
    bool gpk[NLP];
    double pnext[NLP];
    int kpre = 0, pcur = 0;
    if (CORR) {
#pragma unroll
        for (int r = 0; r < NLP; ++r) {
            const int q = FULL ? min(tid + r * NT, LPP - 1) : tid + r * NT;
            const int row = q / PTX, col = q - row * PTX;
            const int gi = i0 + col, gj = j0 + row;
            gpk[r] = FULL || (q < LPP && gi <= g.nx + 2 && gj <= g.ny + 2);
            gpo[r] = gpk[r] ? wrap_p(gi, g.nx) + sy * wrap_p(gj, g.ny) : 0;
            // p of the chunk's first plane, for its first projection
            if (FULL || q < LPP) tp[0][q] = gpk[r] ? cr.p[gpo[r] + sz * k0] : 0.0;
        }
    }
    // interior: 1 <= kp <= nz - 1 is known (no wall rule applies to the plane)
    auto fetch = [&](int kp, bool interior = false) {
        kpre = kp;
        if (CORR) {
            const int ks = interior ? kp : (kp == 0 ? 1 : (kp == g.nz + 1 ? g.nz : kp));
            const bool above = interior || (kp >= 1 && kp <= g.nz - 1);
#pragma unroll
            for (int r = 0; r < NLD; ++r) {
                const long long e0 = gii[r] + sy * gjj[r] + sz * ks;
                pre[0][r] = gk[r] ? qu[e0] : 0.0;
                pre[1][r] = gk[r] ? qv[e0] : 0.0;
                pre[2][r] = gk[r] ? qw[e0] : 0.0;
            }
#pragma unroll
            for (int r = 0; r < NLP; ++r)
                pnext[r] = (above && gpk[r]) ? cr.p[gpo[r] + sz * (kp + 1)] : 0.0;
        } else if (FILL) {
            const int ks = kp == 0 ? 1 : (kp == g.nz + 1 ? g.nz : kp);
#pragma unroll
            for (int f = 0; f < 3; ++f)
#pragma unroll
                for (int r = 0; r < NLD; ++r)
                    pre[f][r] = gk[r] ? Q[f][gii[r] + sy * gjj[r] + sz * ks] : 0.0;
        } else {
#pragma unroll
            for (int f = 0; f < 3; ++f)
#pragma unroll
                for (int r = 0; r < NLD; ++r) pre[f][r] = gk[r] ? Q[f][go[r] + sz * kp] : 0.0;
        }
    };
    // interior: the plane put lies strictly inside (1 <= kp <= nz - 1), so no
    // wall rule of the fill applies (the FULL loop's planes but the last two)
    auto put = [&](int b, bool interior = false) {
        if (CORR) {
            // projection (k_correct) of the plane's source values, then the
            // fill (k_bc_vel) for the wall planes and w's wall row.  tp[pcur]
            // holds p of the source plane ks; p of the plane above goes to the
            // other tile and becomes the source plane's p for the next put
            const int kp = kpre;
            const int ks = interior ? kp : (kp == 0 ? 1 : (kp == g.nz + 1 ? g.nz : kp));
            const bool wall = !interior && (kp == 0 || kp == g.nz + 1);
            const bool above = interior || (kp >= 1 && kp <= g.nz - 1);
            const double dzi = pll::zm(g.dzc_inv, ks), bcu = 0.0 - g.u_crf;
            if (above) {
#pragma unroll
                for (int r = 0; r < NLP; ++r) {
                    const int q = FULL ? min(tid + r * NT, LPP - 1) : tid + r * NT;
                    if (FULL || q < LPP) tp[pcur ^ 1][q] = pnext[r];
                }
            }
            __syncthreads();
            const double *Pc = tp[pcur], *Pn = tp[pcur ^ 1];
#pragma unroll
            for (int r = 0; r < NLD; ++r) {
                const int q = tid + r * NT;
                const int row = q / LTX, col = q - row * LTX;
                const int o = q < LPL ? row * PTX + col : 0;
                const double pc = Pc[o];
                double un = pre[0][r] - (Pc[o + 1] - pc) * cr.dtdxi;
                if (cr.fmask & 1) un = un + shu;
                double vn = pre[1][r] - (Pc[o + PTX] - pc) * cr.dtdyi;
                if (cr.fmask & 2) vn = vn + shv;
                double wn = above ? pre[2][r] - (Pn[o] - pc) * g.dt * dzi : 0.0;
                if (wall) {
                    un = 2.0 * bcu - un;
                    vn = 2.0 * 0.0 - vn;
                    wn = 0.0;
                } else if (!interior && kp == g.nz) {
                    wn = 0.0;
                }
                pre[0][r] = un;
                pre[1][r] = vn;
                pre[2][r] = wn;
            }
            if (above) pcur ^= 1;
        } else if (FILL && !interior) {
            const int kp = kpre;
            if (kp == 0 || kp == g.nz + 1) {
                const double bcu = 0.0 - g.u_crf;
#pragma unroll
                for (int r = 0; r < NLD; ++r) {
                    pre[0][r] = 2.0 * bcu - pre[0][r];
                    pre[1][r] = 2.0 * 0.0 - pre[1][r];
                    pre[2][r] = 0.0;
                }
            } else if (kp == g.nz) {
#pragma unroll
                for (int r = 0; r < NLD; ++r) pre[2][r] = 0.0;
            }
        }
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
            for (int r = 0; r < NLD; ++r) {
                const int q = FULL ? min(tid + r * NT, LPL - 1) : tid + r * NT;
                if (FULL || q < LPL) tl[b][f][q] = pre[f][r];
            }
    };
    fetch(k0);
    put(0);
    if (CORR) __syncthreads();  // put(0)'s pressure reads are done before put(1) writes a tile
    fetch(k0 + 1);
    put(1);
    long long e = i + sy * j + sz * k0;
    double u00m = 0.0, v00m = 0.0, w00m = 0.0, wp0m = 0.0, w0pm = 0.0;
    if (valid) {
        if (FILL) {
            // plane k0 - 1 as the fill leaves it: a wall plane for k0 = 1
            auto fq = [&](const double *q, int f, int ii, int jj) -> double {
                const int kk = k0 - 1;
                if (kk == 0) {
                    if (f == 2) return 0.0;
                    const double bc = f == 0 ? 0.0 - g.u_crf : 0.0;
                    return 2.0 * bc - q[ii + sy * jj + sz * 1];
                }
                return q[ii + sy * jj + sz * kk];  // kk < nz: no wall row here
            };
            u00m = fq(qu, 0, i, j);
            v00m = fq(qv, 1, i, j);
            w00m = fq(qw, 2, i, j);
            wp0m = fq(qw, 2, wrap(i + 1, g.nx), j);
            w0pm = fq(qw, 2, i, wrap(j + 1, g.ny));
        } else if (CORR) {
            u00m = filled(0, i, j, k0 - 1);
            v00m = filled(1, i, j, k0 - 1);
            w00m = filled(2, i, j, k0 - 1);
            wp0m = filled(2, wrap(i + 1, g.nx), j, k0 - 1);
            w0pm = filled(2, i, wrap(j + 1, g.ny), k0 - 1);
        } else {
            u00m = qu[e - sz];
            v00m = qv[e - sz];
            w00m = qw[e - sz];
            wp0m = qw[e - sz + 1];
            w0pm = qw[e - sz + sy];
        }
    }
    __syncthreads();
    double su = 0.0, sv = 0.0;
    const int c0 = (ly + 1) * LTX + lx + 1;  // this cell in a tile
    // plane k from LDS buffer b (plane k+1 in b ^ 1); old: RK2's previous
    // values of plane k when loaded ahead (FULL), else read here
    auto compute = [&](int kv, int b, long long ek, const double *old) {
        // k is the same for the whole workgroup: the z metrics come through
        // the scalar unit (scalar loads wait on their own counter)
        const int k = __builtin_amdgcn_readfirstlane(kv);
        const double *cu = tl[b][0], *cv = tl[b][1], *cw = tl[b][2];
        const double *nu = tl[b ^ 1][0], *nv = tl[b ^ 1][1], *nw = tl[b ^ 1][2];
        Nb n;
        n.u000 = cu[c0]; n.up00 = cu[c0 + 1]; n.um00 = cu[c0 - 1];
        n.u0p0 = cu[c0 + LTX]; n.u0m0 = cu[c0 - LTX]; n.ump0 = cu[c0 - 1 + LTX];
        n.u00p = nu[c0]; n.um0p = nu[c0 - 1]; n.u00m = u00m;
        n.v000 = cv[c0]; n.vp00 = cv[c0 + 1]; n.vm00 = cv[c0 - 1];
        n.v0p0 = cv[c0 + LTX]; n.v0m0 = cv[c0 - LTX]; n.vpm0 = cv[c0 + 1 - LTX];
        n.v00p = nv[c0]; n.v0mp = nv[c0 - LTX]; n.v00m = v00m;
        n.w000 = cw[c0]; n.wp00 = cw[c0 + 1]; n.wm00 = cw[c0 - 1];
        n.w0p0 = cw[c0 + LTX]; n.w0m0 = cw[c0 - LTX]; n.w00p = nw[c0];
        n.w00m = w00m; n.wp0m = wp0m; n.w0pm = w0pm;
        double fu, fv, fw;
        mom_rhs(g, n, k, fu, fv, fw);
        if (BODY) {
            fu = fu + fb.u[ek];
            fv = fv + fb.v[ek];
            fw = fw + fb.w[ek];
        }
        if (SECOND) {
            double ou0, ov0, ow0;
            if constexpr (PROJO) {
                // the expressions of proj() on o
                const double pc = cr.p[ek];
                ou0 = ou[ek] - (cr.p[ek + 1] - pc) * cr.dtdxi;
                if (cr.fmask & 1) ou0 = ou0 + shu;
                ov0 = ov[ek] - (cr.p[ek + sy] - pc) * cr.dtdyi;
                if (cr.fmask & 2) ov0 = ov0 + shv;
                ow0 = k == g.nz ? 0.0 : ow[ek] - (cr.p[ek + sz] - pc) * g.dt * g.dzc_inv[k];
            } else {
                ou0 = old ? old[0] : ou[ek];
                ov0 = old ? old[1] : ov[ek];
                ow0 = old ? old[2] : ow[ek];
            }
            const double un = (n.u000 + g.dt * fu + ou0) * 0.5;
            const double vn = (n.v000 + g.dt * fv + ov0) * 0.5;
            ou[ek] = un;
            ov[ek] = vn;
            ow[ek] = (n.w000 + g.dt * fw + ow0) * 0.5;
            if (MEAN) {
                // both sums are formed (no branch in the loop); a component
                // not in fmask reports 0
                const double wk = pll::zm(g.dzflzi, k);
                su = su + un * wk;
                sv = sv + vn * wk;
            }
        } else {
            ou[ek] = n.u000 + g.dt * fu;
            ov[ek] = n.v000 + g.dt * fv;
            ow[ek] = n.w000 + g.dt * fw;
            if (CORR && cr.nu) {
                cr.nu[ek] = n.u000;
                cr.nv[ek] = n.v000;
                cr.nw[ek] = n.w000;
            }
        }
        u00m = n.u000;
        v00m = n.v000;
        w00m = n.w000;
        wp0m = n.wp00;
        w0pm = n.w0p0;
    };
    if constexpr (FULL) {
        double oldc[3] = {0.0, 0.0, 0.0};
        // every load of the prologue has arrived before the loop starts, so that
        // no wait inside the loop (which covers its first iteration too) has
        // to include the loads in flight at that point
        asm volatile("" : : "v"(u00m), "v"(v00m), "v"(w00m), "v"(wp0m), "v"(w0pm));
        auto step = [&](int k, auto INTERIOR) {
            const int b = (k - k0) & 1;
            if (SECOND) {
                // RK2's old values of plane k, issued before the prefetch so
                // that waiting for them leaves the prefetch in flight
                oldc[0] = ou[e];
                oldc[1] = ov[e];
                oldc[2] = ow[e];
            }
            fetch(min(k + 2, k1 + 1), decltype(INTERIOR)::value);
            // keep the prefetch ahead of the arithmetic (the scheduler would
            // otherwise sink part of it next to its use)
            __builtin_amdgcn_sched_barrier(0);
            compute(k, b, e, SECOND ? oldc : nullptr);
            __syncthreads();                      // plane k's buffer is free
            put(b, decltype(INTERIOR)::value);    // it becomes plane k+2 (the last
            __syncthreads();                      // two: a plane not read)
            e += sz;
        };
        // planes whose plane k+2 lies inside take the put without the wall
        // rules: a wall rule in the loop would let the compiler move the w
        // prefetch into its branch, next to its use
        const int kin = min(k1, g.nz - 3);
        int k = k0;
#pragma unroll 1
        for (; k <= kin; ++k) step(k, std::true_type{});
#pragma unroll 1
        for (; k <= k1; ++k) step(k, std::false_type{});
    } else {
        for (int k = k0; k <= k1; ++k, e += sz) {
            const int b = (k - k0) & 1;
            const bool more = k + 2 <= k1 + 1;
            if (more) fetch(k + 2);
            if (valid) compute(k, b, e, nullptr);
            __syncthreads();  // plane k's buffer is free
            if (more) put(b);  // it becomes plane k+2
            __syncthreads();
        }
    }
    if (MEAN) {
        __shared__ double red[2][NT];
        red[0][tid] = su;
        red[1][tid] = sv;
        __syncthreads();
        for (int h = NT / 2; h > 0; h >>= 1) {
            if (tid < h) {
                red[0][tid] += red[0][tid + h];
                red[1][tid] += red[1][tid + h];
            }
            __syncthreads();
        }
        if (tid == 0) {
            mpart[blockIdx.x] = (fmask & 1) ? red[0][0] : 0.0;
            mpart[gridDim.x + blockIdx.x] = (fmask & 2) ? red[1][0] : 0.0;
        }
    }
}

// correctVel (mod_calcVel.f90:428-453) marching over a k chunk (p(k+1) is the
// next step's p(k)).  FORCE: then forceVel's shift q += vel_force - mean for
// the components in fmask (bit 0 u, bit 1 v), with the means in mean[0..1],
// in the same two roundings as correctVel followed by forceVel.
template <bool FORCE>
__global__ void __launch_bounds__(RK_TX *RK_TY) k_correct(G g, Tiles t, const double *__restrict__ p,
                                                          double *__restrict__ u,
                                                          double *__restrict__ v,
                                                          double *__restrict__ w, double dtdxi,
                                                          double dtdyi, int fmask, double vfu,
                                                          double vfv,
                                                          const double *__restrict__ mean)
{
    int tx, ty, kc;
    rk_tile(t, tx, ty, kc);
    const int i = tx * RK_TX + (int)(threadIdx.x % RK_TX) + 1;
    const int j = ty * RK_TY + (int)(threadIdx.x / RK_TX) + 1;
    if (i > g.nx || j > g.ny) return;
    const int k0 = kc * t.kc + 1, k1 = min(k0 + t.kc - 1, g.nz);
    const long long sy = g.sy, sz = g.sz;
    long long e = i + sy * j + sz * k0;
    double fu = 0.0, fv = 0.0;
    if (FORCE) {
        if (fmask & 1) fu = vfu - mean[0];
        if (fmask & 2) fv = vfv - mean[1];
    }
    double pc = p[e];
    for (int k = k0; k <= k1; ++k, e += sz) {
        const double pz = p[e + sz];
        double un = u[e] - (p[e + 1] - pc) * dtdxi;
        double vn = v[e] - (p[e + sy] - pc) * dtdyi;
        if (FORCE && (fmask & 1)) un = un + fu;
        if (FORCE && (fmask & 2)) vn = vn + fv;
        u[e] = un;
        v[e] = vn;
        w[e] = w[e] - (pz - pc) * g.dt * g.dzc_inv[k];
        pc = pz;
    }
}

#define CELL_IDX                                                   \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;       \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;       \
    const int k = blockIdx.z + 1;                                  \
    if (i > g.nx || j > g.ny) return;                              \
    const long long e = i + g.sy * j + g.sz * k;

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
// KIND 0: |div u| (calcMaxDiv, mod_utils.f90:286-288);
// KIND 1: the CFL sum before the factor dt (calcMaxCFL, mod_utils.f90:224-226)
//   |(u_i + u_{i-1})/2| dx_inv + |(v_j + v_{j-1})/2| dy_inv + |(w_k + w_{k-1})/2| dzf_inv(k).
// NaN cells never win (every comparison with NaN is false), as in the
// reference's `if (x > max)` loops.
template <int KIND>
__global__ void __launch_bounds__(256) k_maxloc_part(G g, const double *__restrict__ u,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ w,
                                                      double *__restrict__ pval,
                                                      long long *__restrict__ pidx)
{
    __shared__ double sv[256];
    __shared__ long long si[256];
    // Rows go to the 8 XCDs (which take workgroups round-robin) as y slabs:
    // XCD x works through rows j in [x ny/8, (x+1) ny/8) of one plane after
    // the other, so the rows j-1 and k-1 that a row reads were fetched into
    // the same L2 shortly before (ny % 8 == 0; plain row order otherwise).
    // The partials are reduced with ties to the lowest cell index, so the
    // result does not depend on the order.
    int j, k;
    if (g.ny % 8 == 0) {
        const int sl = g.ny >> 3, p = (int)(blockIdx.x >> 3);
        j = (int)(blockIdx.x & 7) * sl + p % sl + 1;
        k = p / sl + 1;
    } else {
        j = blockIdx.x % g.ny + 1;
        k = blockIdx.x / g.ny + 1;
    }
    const long long base = g.sy * j + g.sz * k;
    double best = -1.0;
    long long bi = 0;
    for (int i = threadIdx.x + 1; i <= g.nx; i += blockDim.x) {
        const long long e = base + i;
        double d;
        if (KIND == 0)
            d = fabs((u[e] - u[e - 1]) * g.dx_inv + (v[e] - v[e - g.sy]) * g.dy_inv +
                     (w[e] - w[e - g.sz]) * g.dzf_inv[k]);
        else
            d = fabs((u[e] + u[e - 1]) * 0.5) * g.dx_inv + fabs((v[e] + v[e - g.sy]) * 0.5) * g.dy_inv +
                fabs((w[e] + w[e - g.sz]) * 0.5) * g.dzf_inv[k];
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
    // eight partials in flight per thread (one at a time leaves the loop
    // waiting on memory latency: 54 us for 98304 partials)
    const int nb = blockDim.x;
    int t = threadIdx.x;
    for (; t + 7 * nb < n; t += 8 * nb) {
        double o[8];
        long long oi[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            o[q] = pval[t + q * nb];
            oi[q] = pidx[t + q * nb];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (o[q] > best || (o[q] == best && oi[q] < bi)) {
                best = o[q];
                bi = oi[q];
            }
    }
    for (; t < n; t += nb) {
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
// checkNaN (mod_utils.f90:352-394) over every element of a haloed field:
// flags[block] = 1 if the block's stride of elements holds a NaN.
__global__ void __launch_bounds__(256) k_nan_part(const double *__restrict__ q, long long n,
                                                  long long *__restrict__ flags)
{
    int bad = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        bad |= q[t] != q[t];
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) flags[blockIdx.x] = bad;
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : -3; }

Tiles rk_tiles(const G &g, int ty = RK_TY)
{
    Tiles t;
    t.kc = g.kc > 0 ? g.kc : 64;
    t.ntx = (g.nx + RK_TX - 1) / RK_TX;
    t.nty = (g.ny + ty - 1) / ty;
    t.nkc = (g.nz + t.kc - 1) / t.kc;
    t.xcd = ((long long)t.ntx * t.nty * t.nkc) % 8 == 0;
    return t;
}

}  // namespace

extern "C" int pll_rk_blocks(const pll_grid_dev *gd)
{
    const Tiles t = rk_tiles(make_g(gd));
    return t.ntx * t.nty * t.nkc;
}

// runtime flags -> template instantiation
template <typename F>
int with_flags(bool a, bool b, bool c, F &&f)
{
#define PLL_B3(A, B, C) \
    if (a == A && b == B && c == C) return f(std::integral_constant<bool, A>{}, std::integral_constant<bool, B>{}, std::integral_constant<bool, C>{});
    PLL_B3(false, false, false) PLL_B3(false, false, true) PLL_B3(false, true, false)
    PLL_B3(false, true, true) PLL_B3(true, false, false) PLL_B3(true, false, true)
    PLL_B3(true, true, false) PLL_B3(true, true, true)
#undef PLL_B3
    return -1;
}

// Grids whose RK tiles all lie inside take the FULL kernels for RK1 and the
// projecting RK1 (768x384x256, same-process A/B: projecting RK1 1.698 ->
// 1.542 ms, RK1 unchanged); RK2's FULL form needs 145 VGPRs (3 waves per SIMD
// against 4) and runs 1.314 ms against 1.304, so RK2 keeps the general form
// unless PLL_RK2_FULL=1.  PLL_RK_FULL=0 takes the general kernels everywhere.
// (read at every launch, so that one process can time both forms)
bool full_tiles(const G &g, int ty, bool second = false)
{
    const char *e = getenv("PLL_RK_FULL");
    const char *e2 = getenv("PLL_RK2_FULL");
    const bool allow = !(e && atoi(e) == 0), allow2 = e2 && atoi(e2) != 0;
    return allow && (!second || allow2) && g.nx % RK_TX == 0 && g.ny % ty == 0;
}

extern "C" int pll_launch_rk1(const pll_grid_dev *gd, const double *u, const double *v,
                              const double *w, double *u1, double *v1, double *w1,
                              const double *const *body, hipStream_t s)
{
    const G g = make_g(gd);
    const Tiles t = rk_tiles(g);
    const dim3 grid((unsigned)(t.ntx * t.nty * t.nkc)), blk(RK_TX * RK_TY);
    const Body fb = body ? Body{body[0], body[1], body[2]} : Body{nullptr, nullptr, nullptr};
    return with_flags(false, body != nullptr, full_tiles(g, RK_TY), [&](auto, auto B, auto F) {
        hipLaunchKernelGGL((k_rk<false, false, decltype(B)::value, false, RK_TY, false,
                                 decltype(F)::value>),
                           grid, blk, 0, s, g, t, u, v, w, u1, v1, w1, nullptr, 0, fb, Corr{});
        return ok();
    });
}

extern "C" int pll_launch_rk1_corr(const pll_grid_dev *gd, const double *u, const double *v,
                                   const double *w, const double *p, double *nu, double *nv,
                                   double *nw, double *u1, double *v1, double *w1, int fmask,
                                   double vfu, double vfv, const double *mean, hipStream_t s)
{
    const G g = make_g(gd);
    const Tiles t = rk_tiles(g, RK_TY_CORR);
    const dim3 grid((unsigned)(t.ntx * t.nty * t.nkc)), blk(RK_TX * RK_TY_CORR);
    Corr cr;
    cr.p = p;
    cr.nu = nu;
    cr.nv = nv;
    cr.nw = nw;
    cr.dtdxi = g.dt * g.dx_inv;  // as pll_launch_correct_vel
    cr.dtdyi = g.dt * g.dy_inv;
    cr.fmask = mean ? fmask : 0;
    cr.vfu = vfu;
    cr.vfv = vfv;
    cr.mean = mean;
    if (full_tiles(g, RK_TY_CORR))
        hipLaunchKernelGGL((k_rk<false, false, false, true, RK_TY_CORR, false, true>), grid, blk, 0,
                           s, g, t, u, v, w, u1, v1, w1, nullptr, 0,
                           Body{nullptr, nullptr, nullptr}, cr);
    else
        hipLaunchKernelGGL((k_rk<false, false, false, true, RK_TY_CORR, false, false>), grid, blk,
                           0, s, g, t, u, v, w, u1, v1, w1, nullptr, 0,
                           Body{nullptr, nullptr, nullptr}, cr);
    return ok();
}

extern "C" int pll_launch_rk2(const pll_grid_dev *gd, double *u, double *v, double *w,
                              const double *u1, const double *v1, const double *w1,
                              double *mpart, int fmask, const double *const *body, int fill,
                              hipStream_t s)
{
    const G g = make_g(gd);
    const Tiles t = rk_tiles(g);
    const dim3 grid((unsigned)(t.ntx * t.nty * t.nkc)), blk(RK_TX * RK_TY);
    const Body fb = body ? Body{body[0], body[1], body[2]} : Body{nullptr, nullptr, nullptr};
    const bool mean = mpart && fmask;
    return with_flags(mean, body != nullptr, fill != 0, [&](auto M, auto B, auto F) {
        if (full_tiles(g, RK_TY, true))
            hipLaunchKernelGGL((k_rk<true, decltype(M)::value, decltype(B)::value, false, RK_TY,
                                     decltype(F)::value, true>),
                               grid, blk, 0, s, g, t, u1, v1, w1, u, v, w, mean ? mpart : nullptr,
                               mean ? fmask : 0, fb, Corr{});
        else
            hipLaunchKernelGGL((k_rk<true, decltype(M)::value, decltype(B)::value, false, RK_TY,
                                     decltype(F)::value, false>),
                               grid, blk, 0, s, g, t, u1, v1, w1, u, v, w, mean ? mpart : nullptr,
                               mean ? fmask : 0, fb, Corr{});
        return ok();
    });
}

// RK2 of a multi-step run, in place: o = (u, v, w) is the previous step's
// unprojected RK2 output, projected at each cell as it is read (PROJO), and
// is overwritten with the new velocity; q = the RK1 stage; the boundary values
// of the RK1 stage are read from its interior (FILL).
extern "C" int pll_launch_rk2_proj(const pll_grid_dev *gd, double *u, double *v, double *w,
                                   const double *u1, const double *v1, const double *w1,
                                   const double *p, double *mpart, int fmask, double vfu,
                                   double vfv, const double *mean, hipStream_t s)
{
    const G g = make_g(gd);
    const Tiles t = rk_tiles(g);
    const dim3 grid((unsigned)(t.ntx * t.nty * t.nkc)), blk(RK_TX * RK_TY);
    Corr cr;
    cr.p = p;
    cr.nu = cr.nv = cr.nw = nullptr;
    cr.dtdxi = g.dt * g.dx_inv;  // as pll_launch_correct_vel
    cr.dtdyi = g.dt * g.dy_inv;
    cr.fmask = mean ? fmask : 0;
    cr.vfu = vfu;
    cr.vfv = vfv;
    cr.mean = mean;
    const Body fb{nullptr, nullptr, nullptr};
    if (mpart && fmask)
        hipLaunchKernelGGL((k_rk<true, true, false, false, RK_TY, true, false, true>), grid, blk, 0,
                           s, g, t, u1, v1, w1, u, v, w, mpart, fmask, fb, cr);
    else
        hipLaunchKernelGGL((k_rk<true, false, false, false, RK_TY, true, false, true>), grid, blk,
                           0, s, g, t, u1, v1, w1, u, v, w, nullptr, 0, fb, cr);
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
                                      double *v, double *w, int fmask, double vfu, double vfv,
                                      const double *mean, hipStream_t s)
{
    const G g = make_g(gd);
    const Tiles t = rk_tiles(g);
    const double dtdxi = g.dt * g.dx_inv, dtdyi = g.dt * g.dy_inv;  // mod_calcVel.f90:436-437
    const dim3 grid((unsigned)(t.ntx * t.nty * t.nkc));
    const bool force = mean && fmask;
    return with_flags(force, false, false, [&](auto Fo, auto, auto) {
        hipLaunchKernelGGL(k_correct<decltype(Fo)::value>, grid, dim3(RK_TX * RK_TY), 0, s, g, t, p,
                           u, v, w, dtdxi, dtdyi, force ? fmask : 0, vfu, vfv,
                           force ? mean : nullptr);
        return ok();
    });
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
    hipLaunchKernelGGL(k_maxloc_part<0>, dim3(g.ny * g.nz), dim3(256), 0, s, g, u, v, w, pval, pidx);
    return ok();
}

extern "C" int pll_launch_maxcfl_partials(const pll_grid_dev *gd, const double *u, const double *v,
                                          const double *w, double *pval, long long *pidx,
                                          hipStream_t s)
{
    const G g = make_g(gd);
    hipLaunchKernelGGL(k_maxloc_part<1>, dim3(g.ny * g.nz), dim3(256), 0, s, g, u, v, w, pval, pidx);
    return ok();
}

extern "C" int pll_launch_nan_partials(const pll_grid_dev *gd, const double *q, long long *flags,
                                       int nblocks, hipStream_t s)
{
    const G g = make_g(gd);
    const long long n = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    hipLaunchKernelGGL(k_nan_part, dim3(nblocks), dim3(256), 0, s, q, n, flags);
    return ok();
}

extern "C" int pll_launch_maxdiv_final(int n, const double *pval, const long long *pidx,
                                       double *oval, long long *oidx, hipStream_t s)
{
    hipLaunchKernelGGL(k_maxdiv_final, dim3(1), dim3(1024), 0, s, n, pval, pidx, oval, oidx);
    return ok();
}
