/*
 * Host side of the GPU pressure Poisson solver (C, drives the HIP kernels).
 *
 * Mirrors src/src_c/mod_poisson_solver.c of the reference:
 *   init_poisson_solver  :178-309  -> pll_ppe_create + the global instance
 *   execute_poisson_solver :468-605 -> pll_ppe_solve (5 kernels)
 *   free_poisson_solver  :607-632  -> pll_ppe_destroy
 * The eigenvalues follow get_eigen_values (src/src_c/mod_fft_fftw.inc:175-198)
 * and the z coefficients set_trid_coeff (mod_poisson_solver.c:40-103), both
 * evaluated here with the reference's formulas.  The solver state stays on
 * the GPU for the lifetime of the handle.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/powerllel.h"
#include "pll_internal.h"

/* ------------------------------------------------------------------ errors */
static __thread char g_err[512];

const char *pll_last_error(void) { return g_err; }

int pll_set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return pll_set_error(PLL_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

/* ------------------------------------------------------------------ plans */
/* Factor L = 2^a 3^b 5^c 7^d 11^e 13^f into radix stages from
 * {8,6,4,3,2,5,7,11,13}.  E (points per thread) is the first of
 * {12,8,24,6,4,3,5,7,2} that divides L, keeps L/E threads per row within the
 * kernel's limit and divides every radix of the plan, so that all stages run
 * from registers (E = 24 needs more than 256 VGPRs per lane, so it is the
 * last resort of the large ones).  Lengths that no such E fits (mixed
 * prime factors, such as 2^a 5 or 2^a 3 7) take the first E that divides L;
 * their stages whose radix does not divide E run through LDS. */
static int plan_for_E(int L, int E, const int cnt[4], const int cnt_big[2], int reg_only,
                      pll_fftplan *p)
{
    int a = cnt[0], b = cnt[1], n = 0, prod = 1;
    int rad[32];
    while (a >= 3 && E % 8 == 0) { rad[n++] = 8; a -= 3; }
    while (a >= 2 && E % 4 == 0) { rad[n++] = 4; a -= 2; }
    if (a == 1 && b >= 1 && E % 6 == 0) { rad[n++] = 6; a = 0; b--; }
    while (a >= 1 && E % 2 == 0) { rad[n++] = 2; a -= 1; }
    while (b >= 1 && E % 3 == 0) { rad[n++] = 3; b -= 1; }
    /* the rest through LDS */
    while (a >= 3) { rad[n++] = 8; a -= 3; }
    while (a >= 1) { rad[n++] = a >= 2 ? 4 : 2; a -= a >= 2 ? 2 : 1; }
    while (b >= 1) { rad[n++] = 3; b -= 1; }
    for (int i = 0; i < cnt[2]; i++) rad[n++] = 5;
    for (int i = 0; i < cnt[3]; i++) rad[n++] = 7;
    for (int i = 0; i < cnt_big[0]; i++) rad[n++] = 11;
    for (int i = 0; i < cnt_big[1]; i++) rad[n++] = 13;
    if (n > PLL_MAX_STAGES) return 0;
    for (int s = 0; s < n; s++) {
        prod *= rad[s];
        if (reg_only && E % rad[s]) return 0;
    }
    if (prod != L) return 0;
    p->L = L;
    p->E = E;
    p->T = L / E;
    p->nst = n;
    memcpy(p->rad, rad, sizeof(int) * (size_t)n);
    return 1;
}

static int make_plan(int L, pll_fftplan *p)
{
    memset(p, 0, sizeof *p);
    /* PLL_FFT_E selects among compile-time plans of the same length (tuning) */
    const char *pe = getenv("PLL_FFT_E");
    if (pll_fft_ct_plan(L, pe ? atoi(pe) : 0, p)) return PLL_OK;
    memset(p, 0, sizeof *p);
    if (L < 2) return pll_set_error(PLL_EUNSUP, "FFT length %d < 2", L);
    static const int primes[6] = {2, 3, 5, 7, 11, 13};
    int c6[6] = {0}, m = L;
    for (int i = 0; i < 6; i++)
        while (m % primes[i] == 0) { m /= primes[i]; c6[i]++; }
    if (m != 1)
        return pll_set_error(PLL_EUNSUP,
                             "FFT length %d has a prime factor above 13 (radices 2, 3, 5, 7, "
                             "11 and 13 are supported)", L);
    const int cnt[4] = {c6[0], c6[1], c6[2], c6[3]}, big[2] = {c6[4], c6[5]};
    static const int Es[] = {12, 8, 24, 6, 4, 3, 5, 7, 2};
    for (int reg_only = 1; reg_only >= 0; reg_only--)
        for (size_t ie = 0; ie < sizeof Es / sizeof Es[0]; ie++) {
            const int E = Es[ie];
            if (L % E != 0 || !pll_fft_supported_E(E) || L / E > pll_fft_max_threads(E)) continue;
            if (plan_for_E(L, E, cnt, big, reg_only, p)) return PLL_OK;
        }
    return pll_set_error(PLL_EUNSUP, "no FFT plan for length %d", L);
}

/* exp(-2 pi i t / n), t in [0, n), as interleaved (re, im) doubles */
static double *twiddles(int n)
{
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    if (!t) return NULL;
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int i = 0; i < n; i++) {
        long double ang = twopi * (long double)i / (long double)n;
        t[2 * i] = (double)cosl(ang);
        t[2 * i + 1] = (double)(-sinl(ang));
    }
    return t;
}

/* get_eigen_values, mod_fft_fftw.inc:175-183, periodic case, 1-based i */
static void eigen_pp(int n, double h, double *lam)
{
    const double pi = acos(-1.0);
    for (int i = 1; i <= n; i++) lam[i - 1] = 2.0 * (cos(2.0 * pi * (i - 1.0) / n) - 1.0);
    for (int i = 0; i < n; i++) lam[i] /= (h * h); /* mod_poisson_solver.c:231-237 */
}

/* set_trid_coeff, mod_poisson_solver.c:40-103, 'c' grid, walls at both ends.
 * dzf: n+2 entries with ghosts. */
static int trid_coeff(int n, const double *dzf, const char *bct, double *a, double *b,
                      double *c)
{
    for (int k = 0; k < n; k++) {
        a[k] = 2.0 / (dzf[k + 1] * (dzf[k] + dzf[k + 1]));
        c[k] = 2.0 / (dzf[k + 1] * (dzf[k + 2] + dzf[k + 1]));
    }
    for (int k = 0; k < n; k++) b[k] = -a[k] - c[k];
    for (int side = 0; side < 2; side++) {
        char bc = bct[side];
        double factor;
        if (bc == 'D') factor = -1.0;
        else if (bc == 'N') factor = 1.0;
        else return pll_set_error(PLL_EUNSUP, "z boundary type '%c' (want N or D)", bc);
        if (side == 0) {
            b[0] += factor * a[0];
            a[0] *= fabs(factor) - 1.0;
        } else {
            b[n - 1] += factor * c[n - 1];
            c[n - 1] *= fabs(factor) - 1.0;
        }
    }
    return PLL_OK;
}

/* a + b + c to within one rounding of the exact sum (Knuth's TwoSum on both
 * additions).  The row sums of the z operator are zero up to the rounding of
 * b = -a - c, and that rounding is part of the operator the solve inverts, so
 * the residual of the refined form needs it exactly. */
static double rowsum3(double a, double b, double c)
{
    const double s = a + b, bv = s - a, e1 = (a - (s - bv)) + (b - bv);
    const double t = s + c, cv = t - s, e2 = (s - (t - cv)) + (c - cv);
    return t + (e1 + e2);
}

/* ------------------------------------------------------------------ handle */
#define PLL_TIMING_RING 128
#define PLL_HOST_SLOTS 3
struct pll_ppe {
    int nx, ny, nz, L;
    pll_ppe_dev d;
    void *W, *tw_x, *tw_y, *a, *c, *piv;
    void *bz, *lamx, *lamy, *order, *order32;  /* partitioned z solve */
    void *tw_x32, *tw_y32, *W32, *P64, *r32; /* refined form (allocated on first use) */
    void *gauge;                      /* top-plane partial sums + mean (refined form) */
    void *low_rx, *low_y, *low_x;     /* lowest modes in FP64 (refined form) */
    int sweeps;                       /* pll_ppe_solve: -1 FP64, >= 0 FP32 + sweeps */
    double tol;                       /* > 0: pll_ppe_solve refines to this tolerance,
                                         with at most `sweeps` sweeps */
    void *nrm;                        /* tolerance mode: max-norm partials (2 + 2 ny nz) */
    int last_sweeps;                  /* tolerance mode: sweeps and error estimate of */
    double last_est;                  /* the last solve                              */
    double *stage;   /* haloed staging buffer for host-pointer calls */
    size_t stage_bytes;
    /* host-pointer calls: pinned staging ring, copy/compute streams, events,
     * the host copy threads (all made by the first host-pointer call) */
    void *pin[PLL_HOST_SLOTS];
    size_t slot_planes;
    hipStream_t hs_copy, hs_comp;
    hipEvent_t ev_up[PLL_HOST_SLOTS], ev_x[PLL_HOST_SLOTS], ev_dn[PLL_HOST_SLOTS];
    int host_ready;
    pll_copy_pool *pool;
    size_t dev_bytes;
    int timing;
    int nrec;                             /* solves recorded since timing on */
    hipEvent_t ev[PLL_TIMING_RING][6];    /* per-solve stage boundaries      */
};

static int dmalloc(pll_ppe *h, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = NULL;
        return pll_set_error(PLL_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    h->dev_bytes += bytes;
    return PLL_OK;
}

int pll_ppe_destroy(pll_ppe *h)
{
    if (!h) return PLL_OK;
    void *bufs[] = {h->W, h->tw_x, h->tw_y, h->a, h->c, h->piv, h->stage,
                    h->bz, h->lamx, h->lamy, h->order, h->tw_x32, h->tw_y32,
                    h->W32, h->P64, h->r32, h->gauge, h->nrm, h->low_rx, h->low_y,
                    h->low_x, h->order32};
    for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; i++)
        if (bufs[i]) hipFree(bufs[i]);
    if (h->timing)
        for (int r = 0; r < PLL_TIMING_RING; r++)
            for (int i = 0; i < 6; i++) hipEventDestroy(h->ev[r][i]);
    if (h->host_ready) {
        hipStreamSynchronize(h->hs_copy);
        hipStreamSynchronize(h->hs_comp);
        for (int i = 0; i < PLL_HOST_SLOTS; i++) {
            hipEventDestroy(h->ev_up[i]);
            hipEventDestroy(h->ev_x[i]);
            hipEventDestroy(h->ev_dn[i]);
        }
        hipStreamDestroy(h->hs_copy);
        hipStreamDestroy(h->hs_comp);
    }
    for (int i = 0; i < PLL_HOST_SLOTS; i++)
        if (h->pin[i]) hipHostFree(h->pin[i]);
    pll_copy_pool_destroy(h->pool);
    free(h);
    return PLL_OK;
}

static int choose_launch(pll_ppe *h)
{
    /* x kernels: rows per workgroup, <= max threads and <= 64 KiB LDS */
    const pll_fftplan *px = &h->d.px, *py = &h->d.py;
    int maxt = pll_fft_max_threads(px->E);
    int rw = 256 / px->T;
    {
        const char *pr = getenv("PLL_X_ROWS"); /* tuning override of the rows per workgroup */
        if (pr && atoi(pr) > 0) rw = atoi(pr);
    }
    if (rw < 1) rw = 1;
    while (rw > 1 && (rw * px->T > maxt || (size_t)rw * (px->L + 1) * 16 > 65536)) rw--;
    if (rw * px->T > maxt || (size_t)rw * (px->L + 1) * 16 > 65536)
        return pll_set_error(PLL_EUNSUP, "nx=%d too large for the x transform kernel", h->nx);
    h->d.x_rows_per_wg = rw;
    /* y kernels: adjacent columns per workgroup (8 gives 128-byte FP64 row
     * segments; XCD-contiguous tile numbering lets narrower tiles share lines
     * in L2), at most 80 KB of LDS so that two workgroups fit a CU */
    maxt = pll_fft_max_threads(py->E);
    int cols = 8;
    {
        const char *pc = getenv("PLL_Y_COLS"); /* tuning override of the y tile width */
        if (pc && atoi(pc) > 0) cols = atoi(pc);
    }
    if (cols > h->L) cols = h->L;
    while (cols > 1 && (cols * py->T > maxt || (size_t)cols * (py->L + 1) * 16 > 81920)) cols--;
    if (cols * py->T > maxt || (size_t)cols * (py->L + 1) * 16 > 81920)
        return pll_set_error(PLL_EUNSUP, "ny=%d too large for the y transform kernel", h->ny);
    h->d.y_cols_per_wg = cols;
    /* FP32 y kernels: 16 complex columns make 128-byte row segments (8 would
     * fetch each line twice); their workgroups may hold 512 threads */
    int cols32 = 16;
    {
        const char *pc = getenv("PLL_Y_COLS32"); /* tuning override */
        if (pc && atoi(pc) > 0) cols32 = atoi(pc);
    }
    if (cols32 > h->L) cols32 = h->L;
    maxt = pll_fft_max_threads32(py->E);
    while (cols32 > 1 && (cols32 * py->T > maxt || (size_t)cols32 * (py->L + 1) * 8 > 81920)) cols32--;
    h->d.y_cols32_per_wg = cols32 < 1 ? 1 : cols32;
    /* x kernels of the refined form: workgroups numbered per XCD, so that the
     * rows of neighbouring planes (read again by the residual) share an L2
     * (refined k = 2 at 768x384x256: 2.97 -> 2.89 ms with the 16-column FP32
     * y tiles); the FP64 x kernels run 3-5 us slower that way.
     * PLL_X_XCD: 0 none, 1 FP32 kernels (default), 2 all */
    const char *px_ = getenv("PLL_X_XCD");
    h->d.x_xcd = px_ ? atoi(px_) : 1;
    /* and for the residual's stencil, each XCD in its own y slab of every
     * plane, so that the rows' y and z neighbours are still in its L2.
     * PLL_X_SLAB: 0 none, 1 stencil sources (default), 2 all per-XCD x kernels */
    const char *ps_ = getenv("PLL_X_SLAB");
    h->d.x_slab = ps_ ? atoi(ps_) : 1;
    return PLL_OK;
}

/* Layout of the partitioned z solve: M rows per lane (M in {8,12,16,4}),
 * P = nz/M lanes per column (a power of two, 2..64, so a column's lanes share
 * one wavefront), S columns per tile (16, 8 or 4 doubles = one 128/64/32-byte
 * row segment); LDS holds the tile twice (values and, for low-mode tiles,
 * pivots).  Returns 0 (streaming solve) if nothing fits. */
static int part_fits(int nx, int nz, int m, int *M, int *P, int *S)
{
    /* widest first: a tile row of 32 FP64 columns is 256 contiguous bytes,
     * which HBM serves faster than 128-byte pieces (768x384x256: 0.270 ->
     * 0.253 ms against 16 columns, same box) */
    static const int Ss[] = {32, 16, 8, 4};
    if (m <= 0 || nz % m || !pll_trid_part_supported(m)) return 0;
    const int p = nz / m;
    if (p < 2 || p > 64 || (p & (p - 1))) return 0;
    const char *ps = getenv("PLL_TRID_S"); /* tuning override of the tile width */
    for (size_t j = 0; j < sizeof Ss / sizeof Ss[0]; j++) {
        const int sc = Ss[j];
        if (ps && atoi(ps) != sc) continue;
        if (nx % sc || sc * p > 512 || (size_t)nz * sc * 8 + (size_t)nz * 24 + (size_t)p * 8 > 163840)
            continue;
        *M = m;
        *P = p;
        *S = sc;
        return 1;
    }
    return 0;
}

static int choose_part(int nx, int nz, int *M, int *P, int *S)
{
    static const int Ms[] = {8, 12, 16, 4};
    const char *pm = getenv("PLL_TRID_M"); /* tuning override of the rows per lane */
    if (pm) return part_fits(nx, nz, atoi(pm), M, P, S);
    /* prefer P = 16 lanes per column (one DPP row) unless PLL_TRID_DPP=0 */
    const char *pd = getenv("PLL_TRID_DPP");
    if (!(pd && atoi(pd) == 0) && nz % 16 == 0 && part_fits(nx, nz, nz / 16, M, P, S)) return 1;
    for (size_t i = 0; i < sizeof Ms / sizeof Ms[0]; i++)
        if (part_fits(nx, nz, Ms[i], M, P, S)) return 1;
    return 0;
}

/* Launch order of the z-solve tiles (SW columns each, nx/SW per y row): tiles
 * with exact-path columns (q0 < qx, iy < iyx) first; workgroup b runs tile
 * order[b].  Workgroups reach the 8 XCDs round-robin, so each XCD is handed a
 * contiguous run of the list: neighbouring tiles (sharing 128-byte lines in
 * FP32 storage) then meet in one L2 (PLL_TRID_XCD=0 keeps the plain order).
 * Returns a malloc'd array, NULL when out of memory. */
static int *tile_order(int nx, int ny, int SW, int qx, int iyx)
{
    const int ntx = nx / SW, ntiles = ntx * ny;
    int *order = (int *)malloc(sizeof(int) * (size_t)ntiles);
    if (!order) return NULL;
    int n = 0;
    for (int pass = 0; pass < 2; pass++)
        for (int t = 0; t < ntiles; t++) {
            const int my = t / ntx, q0 = (t % ntx) * SW;
            const int iy = my <= ny - my ? my : ny - my;
            const int ex = (q0 < qx && iy < iyx);
            if (ex == (pass == 0)) order[n++] = t;
        }
    const char *px = getenv("PLL_TRID_XCD");
    if (ntiles % 8 == 0 && !(px && atoi(px) == 0)) {
        int *perm = (int *)malloc(sizeof(int) * (size_t)ntiles);
        if (!perm) {
            free(order);
            return NULL;
        }
        for (int b = 0; b < ntiles; b++) perm[b] = order[(b % 8) * (ntiles / 8) + b / 8];
        free(order);
        order = perm;
    }
    return order;
}

/* Eigenvalue magnitude below which a column is solved with the sequential
 * recurrence (relative to the largest diagonal of the z operator): the block
 * decomposition's error grows like the condition number of the column. */
#define PLL_EXACT_REL 2e-5

int pll_ppe_create(pll_ppe **out, int nx, int ny, int nz, double dx, double dy,
                   const double *dzf_global, const char *bctype_x, const char *bctype_y,
                   const char *bctype_z)
{
    if (!out) return pll_set_error(PLL_EINVAL, "out is NULL");
    *out = NULL;
    if (!dzf_global || !bctype_x || !bctype_y || !bctype_z)
        return pll_set_error(PLL_EINVAL, "NULL argument");
    if (strcmp(bctype_x, "PP") != 0 || strcmp(bctype_y, "PP") != 0)
        return pll_set_error(PLL_EUNSUP, "x and y must be periodic (\"PP\"), got \"%s\",\"%s\"",
                             bctype_x, bctype_y);
    if (strlen(bctype_z) != 2)
        return pll_set_error(PLL_EINVAL, "bctype_z must have 2 characters");
    if (nx < 4 || ny < 2 || nz < 2 || nx % 2 || ny % 2)
        return pll_set_error(PLL_EINVAL, "need even nx >= 4, even ny >= 2, nz >= 2 (got %d %d %d)",
                             nx, ny, nz);
    if (!(dx > 0.0) || !(dy > 0.0)) return pll_set_error(PLL_EINVAL, "dx, dy must be > 0");
    for (int k = 0; k < nz + 2; k++)
        if (!(dzf_global[k] > 0.0)) return pll_set_error(PLL_EINVAL, "dzf_global[%d] <= 0", k);

    pll_ppe *h = (pll_ppe *)calloc(1, sizeof *h);
    if (!h) return pll_set_error(PLL_ENOMEM, "calloc failed");
    h->nx = nx;
    h->ny = ny;
    h->nz = nz;
    h->L = nx / 2;
    int rc;
    if ((rc = make_plan(h->L, &h->d.px)) || (rc = make_plan(ny, &h->d.py)) ||
        (rc = choose_launch(h))) {
        free(h);
        return rc;
    }
    h->d.nx = nx;
    h->d.ny = ny;
    h->d.nz = nz;
    h->d.L = h->L;
    h->d.normfactor = 1.0 / (1.0 * nx * 1.0 * ny); /* mod_fft_fftw.inc:154 */
    h->d.pin_gauge = (bctype_z[0] == 'N' && bctype_z[1] == 'N');

    const int L = h->L, nyh = ny / 2 + 1;
    const size_t N = (size_t)nx * ny * nz;
    const size_t npiv = (size_t)nz * nyh * (L + 1);
    double *a = (double *)malloc(sizeof(double) * (size_t)nz);
    double *c = (double *)malloc(sizeof(double) * (size_t)nz);
    /* bz | row sums | transposed off-diagonals aT, cT */
    double *bz = (double *)malloc(sizeof(double) * 4 * (size_t)nz);
    double *lx = (double *)malloc(sizeof(double) * (size_t)nx);
    double *ly = (double *)malloc(sizeof(double) * (size_t)ny);
    double *piv = (double *)malloc(sizeof(double) * npiv);
    double *twx = twiddles(nx), *twy = twiddles(ny);
    if (!a || !c || !bz || !lx || !ly || !piv || !twx || !twy) {
        rc = pll_set_error(PLL_ENOMEM, "host allocation failed");
        goto fail;
    }
    if ((rc = trid_coeff(nz, dzf_global, bctype_z, a, bz, c))) goto fail;
    for (int k = 0; k < nz; k++) bz[nz + k] = rowsum3(a[k], bz[k], c[k]);
    /* transposed operator: sub-diagonal c_{k-1}, super-diagonal a_{k+1}; its
     * pivots equal the forward ones (each uses the product a_k c_{k-1}) */
    for (int k = 0; k < nz; k++) {
        bz[2 * nz + k] = k > 0 ? c[k - 1] : 0.0;
        bz[3 * nz + k] = k < nz - 1 ? a[k + 1] : 0.0;
    }
    eigen_pp(nx, dx, lx);
    eigen_pp(ny, dy, ly);
    /* pivots: b_k = bz_k + lambda; b_k -= (a_k / b_{k-1}) c_{k-1}
     * (mod_poisson_solver.c:257-275) for every distinct lambda */
    for (int iy = 0; iy < nyh; iy++) {
        for (int ix = 0; ix <= L; ix++) {
            const double lam = lx[ix] + ly[iy];
            const size_t col = (size_t)iy * (L + 1) + ix;
            double bprev = bz[0] + lam;
            piv[col] = bprev;
            for (int k = 1; k < nz; k++) {
                double bk = bz[k] + lam;
                double a_tmp = a[k] / bprev;
                bk -= (a_tmp * c[k - 1]);
                piv[(size_t)k * nyh * (L + 1) + col] = bk;
                bprev = bk;
            }
        }
    }
    h->sweeps = -1;
    h->d.idx2 = 1.0 / (dx * dx);
    h->d.idy2 = 1.0 / (dy * dy);
    if ((rc = dmalloc(h, &h->bz, sizeof(double) * 4 * (size_t)nz)) ||
        (rc = dmalloc(h, &h->W, sizeof(double) * N)) ||
        (rc = dmalloc(h, &h->tw_x, sizeof(double) * 2 * (size_t)nx)) ||
        (rc = dmalloc(h, &h->tw_y, sizeof(double) * 2 * (size_t)ny)) ||
        (rc = dmalloc(h, &h->a, sizeof(double) * (size_t)nz)) ||
        (rc = dmalloc(h, &h->c, sizeof(double) * (size_t)nz)) ||
        (rc = dmalloc(h, &h->piv, sizeof(double) * npiv)))
        goto fail;
#define H2D(dst, src, bytes)                                                              \
    do {                                                                                  \
        hipError_t e_ = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);                \
        if (e_ != hipSuccess) {                                                           \
            rc = pll_set_error(PLL_EHIP, "hipMemcpy H2D failed: %s", hipGetErrorString(e_)); \
            goto fail;                                                                    \
        }                                                                                 \
    } while (0)
    H2D(h->tw_x, twx, sizeof(double) * 2 * (size_t)nx);
    H2D(h->tw_y, twy, sizeof(double) * 2 * (size_t)ny);
    H2D(h->a, a, sizeof(double) * (size_t)nz);
    H2D(h->c, c, sizeof(double) * (size_t)nz);
    H2D(h->piv, piv, sizeof(double) * npiv);
    H2D(h->bz, bz, sizeof(double) * 4 * (size_t)nz);
#undef H2D
    h->d.tw_x = h->tw_x;
    h->d.tw_y = h->tw_y;
    h->d.a = (const double *)h->a;
    h->d.c = (const double *)h->c;
    h->d.piv = (const double *)h->piv;
    h->d.bz = (const double *)h->bz;
    h->d.dz = (const double *)h->bz + nz;
    {
        int M, P, S;
        if (!getenv("PLL_TRID_STREAMING") && choose_part(nx, nz, &M, &P, &S)) {
            const int ntx = nx / S, ntiles = ntx * ny;
            /* exact-path thresholds: |lambda| < tau * max|bz| */
            double bmax = 0.0;
            for (int k = 0; k < nz; k++) bmax = fabs(bz[k]) > bmax ? fabs(bz[k]) : bmax;
            const double thr = PLL_EXACT_REL * bmax;
            int iyx = 0;
            while (iyx < nyh && fabs(ly[iyx]) < thr) iyx++;
            int ixmax = -1;
            while (ixmax + 1 <= L && fabs(lx[ixmax + 1]) < thr) ixmax++;
            int qx = iyx > 0 ? 2 * (ixmax + 1) : 0;
            const int qx0 = qx;
            qx = ((qx0 + S - 1) / S) * S;
            int *order = tile_order(nx, ny, S, qx, iyx);
            if (!order) {
                rc = pll_set_error(PLL_ENOMEM, "host allocation failed");
                goto fail;
            }
            /* FP32 storage: tiles of 2 S columns (256-byte row segments at
             * S = 32), if their LDS fits (PLL_TRID32_H=1: S columns) */
            const char *ph = getenv("PLL_TRID32_H");
            const int h32 = (S >= 16 && nx % (2 * S) == 0 &&
                             (size_t)nz * 2 * S * 8 + (size_t)nz * 24 + (size_t)P * 8 <= 163840 &&
                             !(ph && atoi(ph) == 1)) ? 2 : 1;
            const int qx32 = ((qx0 + 2 * S - 1) / (2 * S)) * (2 * S);
            if (h32 == 2) {
                int *order32 = tile_order(nx, ny, 2 * S, qx32, iyx);
                if (!order32 || (rc = dmalloc(h, &h->order32, sizeof(int) * (size_t)(ntiles / 2)))) {
                    free(order32);
                    free(order);
                    if (!rc) rc = pll_set_error(PLL_ENOMEM, "host allocation failed");
                    goto fail;
                }
                const hipError_t e32 = hipMemcpy(h->order32, order32, sizeof(int) * (size_t)(ntiles / 2),
                                                 hipMemcpyHostToDevice);
                free(order32);
                if (e32 != hipSuccess) {
                    free(order);
                    rc = pll_set_error(PLL_EHIP, "copy of z-solve tables failed");
                    goto fail;
                }
            }
            h->d.tp.h32 = h32;
            {
                /* FP64 storage: two 16-column groups make the 256-byte row
                 * segments that 32-column tiles give where P = 32 lanes per
                 * column leave room for 16 columns only (nz = 384); at M = 16
                 * rows per lane the second group's values would spill
                 * (PLL_TRID64_H=1: one group) */
                const char *p64 = getenv("PLL_TRID64_H");
                h->d.tp.h64 =
                    (h32 == 2 && S == 16 && M <= 12 && !(p64 && atoi(p64) == 1)) ? 2 : 1;
            }
            h->d.tp.qx32 = qx32;
            h->d.tp.order32 = (const int *)h->order32;
            double *lxh = (double *)malloc(sizeof(double) * (size_t)(L + 1));
            if (!lxh) {
                free(order);
                rc = pll_set_error(PLL_ENOMEM, "host allocation failed");
                goto fail;
            }
            for (int i = 0; i <= L; i++) lxh[i] = lx[i];
            if ((rc = dmalloc(h, &h->lamx, sizeof(double) * (size_t)(L + 1))) ||
                (rc = dmalloc(h, &h->lamy, sizeof(double) * (size_t)nyh)) ||
                (rc = dmalloc(h, &h->order, sizeof(int) * (size_t)ntiles)) ||
                (hipMemcpy(h->lamx, lxh, sizeof(double) * (size_t)(L + 1), hipMemcpyHostToDevice) != hipSuccess) ||
                (hipMemcpy(h->lamy, ly, sizeof(double) * (size_t)nyh, hipMemcpyHostToDevice) != hipSuccess) ||
                (hipMemcpy(h->order, order, sizeof(int) * (size_t)ntiles, hipMemcpyHostToDevice) != hipSuccess)) {
                free(order);
                free(lxh);
                if (!rc) rc = pll_set_error(PLL_EHIP, "copy of z-solve tables failed");
                goto fail;
            }
            free(order);
            free(lxh);
            h->d.tp.M = M;
            h->d.tp.P = P;
            h->d.tp.S = S;
            {
                const char *pd = getenv("PLL_TRID_DPP");
                h->d.tp.dpp16 = (P == 16) && !(pd && atoi(pd) == 0);
            }
            h->d.tp.qx = qx;
            h->d.tp.iyx = iyx;
            h->d.tp.bz = (const double *)h->bz;
            h->d.tp.lamx = (const double *)h->lamx;
            h->d.tp.lamy = (const double *)h->lamy;
            h->d.tp.order = (const int *)h->order;
        }
    }
    free(a); free(c); free(bz); free(lx); free(ly); free(piv); free(twx); free(twy);
    *out = h;
    return PLL_OK;
fail:
    free(a); free(c); free(bz); free(lx); free(ly); free(piv); free(twx); free(twy);
    pll_ppe_destroy(h);
    return rc;
}

size_t pll_ppe_device_bytes(const pll_ppe *h) { return h ? h->dev_bytes : 0; }

int pll_ppe_set_timing(pll_ppe *h, int on)
{
    if (!h) return pll_set_error(PLL_EINVAL, "NULL handle");
    if (on && !h->timing) {
        for (int r = 0; r < PLL_TIMING_RING; r++)
            for (int i = 0; i < 6; i++) HIPCHK(hipEventCreate(&h->ev[r][i]));
        h->timing = 1;
    } else if (!on && h->timing) {
        for (int r = 0; r < PLL_TIMING_RING; r++)
            for (int i = 0; i < 6; i++) hipEventDestroy(h->ev[r][i]);
        h->timing = 0;
    }
    h->nrec = 0;
    return PLL_OK;
}

int pll_ppe_stage_times(pll_ppe *h, float ms[5])
{
    if (!h || !ms) return pll_set_error(PLL_EINVAL, "NULL argument");
    if (!h->timing || h->nrec == 0) return pll_set_error(PLL_ESTATE, "no timed solve recorded");
    const int r = (h->nrec - 1) % PLL_TIMING_RING;
    HIPCHK(hipEventSynchronize(h->ev[r][5]));
    for (int i = 0; i < 5; i++) HIPCHK(hipEventElapsedTime(&ms[i], h->ev[r][i], h->ev[r][i + 1]));
    return PLL_OK;
}

int pll_ppe_stage_times_mean(pll_ppe *h, double ms[5], int *nsolves)
{
    if (!h || !ms) return pll_set_error(PLL_EINVAL, "NULL argument");
    if (!h->timing || h->nrec == 0) return pll_set_error(PLL_ESTATE, "no timed solve recorded");
    const int n = h->nrec < PLL_TIMING_RING ? h->nrec : PLL_TIMING_RING;
    for (int i = 0; i < 5; i++) ms[i] = 0.0;
    for (int q = 0; q < n; q++) {
        const int r = (h->nrec - 1 - q) % PLL_TIMING_RING;
        HIPCHK(hipEventSynchronize(h->ev[r][5]));
        for (int i = 0; i < 5; i++) {
            float t;
            HIPCHK(hipEventElapsedTime(&t, h->ev[r][i], h->ev[r][i + 1]));
            ms[i] += t;
        }
    }
    for (int i = 0; i < 5; i++) ms[i] /= n;
    if (nsolves) *nsolves = n;
    return PLL_OK;
}

static int launch_rc(int rc, const char *what)
{
    if (rc == 0) return PLL_OK;
    if (rc == -1) return pll_set_error(PLL_EUNSUP, "%s: launch shape out of range", what);
    if (rc == -5) return pll_set_error(PLL_EUNSUP, "%s: no kernel for this FFT size", what);
    return pll_set_error(PLL_EHIP, "%s: kernel launch failed: %s", what,
                         hipGetErrorString(hipGetLastError()));
}

int pll_ppe_solve_refined(pll_ppe *h, double *d_var, int sweeps, void *stream);

/* FP64 solve; the right-hand side is the interior of d_var (q == NULL) or
 * formed from the velocity by the x forward transform, in which case the
 * boundary fill of d_var is written with the solution. */
static int solve64(pll_ppe *h, const pll_div_src *q, double *d_var, hipStream_t s)
{
    int rc;
    hipEvent_t *ev = h->timing ? h->ev[h->nrec % PLL_TIMING_RING] : NULL;
    if (ev) HIPCHK(hipEventRecord(ev[0], s));
    rc = q ? pll_launch_xfwd_div(&h->d, q, h->W, s) : pll_launch_xfwd(&h->d, d_var, h->W, s);
    if ((rc = launch_rc(rc, "x forward"))) return rc;
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
    if ((rc = launch_rc(pll_launch_yfwd(&h->d, h->W, s), "y forward"))) return rc;
    if (ev) HIPCHK(hipEventRecord(ev[2], s));
    if ((rc = launch_rc(pll_launch_trid(&h->d, h->W, s), "z solve"))) return rc;
    if (ev) HIPCHK(hipEventRecord(ev[3], s));
    if ((rc = launch_rc(pll_launch_ybwd(&h->d, h->W, s), "y backward"))) return rc;
    if (ev) HIPCHK(hipEventRecord(ev[4], s));
    rc = q ? pll_launch_xbwd_fill(&h->d, h->W, d_var, s) : pll_launch_xbwd(&h->d, h->W, d_var, s);
    if ((rc = launch_rc(rc, "x backward"))) return rc;
    if (ev) {
        HIPCHK(hipEventRecord(ev[5], s));
        h->nrec++;
    }
    return PLL_OK;
}

int pll_ppe_solve(pll_ppe *h, double *d_var, void *stream)
{
    if (!h) return pll_set_error(PLL_ESTATE, "solver not initialised");
    if (!d_var) return pll_set_error(PLL_EINVAL, "var is NULL");
    if (h->tol > 0.0)
        return pll_ppe_solve_refined_tol(h, d_var, h->tol, h->sweeps, NULL, NULL, stream);
    if (h->sweeps >= 0) return pll_ppe_solve_refined(h, d_var, h->sweeps, stream);
    return solve64(h, NULL, d_var, (hipStream_t)stream);
}

int pll_ppe_solve_div(pll_ppe *h, const pll_div_src *q, double *d_var, void *stream)
{
    if (!h) return pll_set_error(PLL_ESTATE, "solver not initialised");
    if (!q || !q->u || !q->v || !q->w || !q->dzf_inv || !d_var)
        return pll_set_error(PLL_EINVAL, "NULL field");
    if (h->sweeps >= 0)
        return pll_set_error(PLL_ESTATE, "the fused right-hand side is for the FP64 solve only");
    return solve64(h, q, d_var, (hipStream_t)stream);
}

int pll_ppe_refinement(const pll_ppe *h) { return h ? h->sweeps : -1; }

int pll_ppe_solve_adjoint(pll_ppe *h, double *d_var, void *stream)
{
    if (!h) return pll_set_error(PLL_ESTATE, "solver not initialised");
    if (!d_var) return pll_set_error(PLL_EINVAL, "var is NULL");
    hipStream_t s = (hipStream_t)stream;
    /* The x and y transforms and the eigenvalue scaling are orthogonal up to
     * a common factor, so the transpose of the solve is the same pipeline
     * with each z system transposed.  The pinned (0,0) system transposes to
     * the same pinned form (its top row and column drop out). */
    pll_ppe_dev dT = h->d;
    dT.a = (const double *)h->bz + 2 * (size_t)h->nz;
    dT.c = (const double *)h->bz + 3 * (size_t)h->nz;
    int rc;
    if ((rc = launch_rc(pll_launch_xfwd(&dT, d_var, h->W, s), "x forward"))) return rc;
    if ((rc = launch_rc(pll_launch_yfwd(&dT, h->W, s), "y forward"))) return rc;
    if ((rc = launch_rc(pll_launch_trid(&dT, h->W, s), "z solve (transposed)"))) return rc;
    if ((rc = launch_rc(pll_launch_ybwd(&dT, h->W, s), "y backward"))) return rc;
    return launch_rc(pll_launch_xbwd(&dT, h->W, d_var, s), "x backward");
}

/* ------------------------------------------------------------------ refined form */
/* FP32 solve + FP64 residual sweeps.  The spectral array, the FFT arithmetic
 * and the correction solves are FP32 (the z recurrence still runs in FP64
 * registers); f stays the caller's FP64 right-hand side, p is accumulated in
 * an FP64 work array and every sweep solves L c = f - L p for a correction. */
static int ensure_fp32(pll_ppe *h)
{
    if (h->W32) return PLL_OK;
    const size_t N = (size_t)h->nx * h->ny * h->nz;
    int rc;
    float *tx = (float *)malloc(sizeof(float) * 2 * (size_t)h->nx);
    float *ty = (float *)malloc(sizeof(float) * 2 * (size_t)h->ny);
    double *dx64 = twiddles(h->nx), *dy64 = twiddles(h->ny);
    if (!tx || !ty || !dx64 || !dy64) {
        free(tx); free(ty); free(dx64); free(dy64);
        return pll_set_error(PLL_ENOMEM, "host allocation failed");
    }
    for (int i = 0; i < 2 * h->nx; i++) tx[i] = (float)dx64[i];
    for (int i = 0; i < 2 * h->ny; i++) ty[i] = (float)dy64[i];
    free(dx64);
    free(dy64);
    /* a failed earlier attempt may have left some of these allocated */
    if ((!h->tw_x32 && (rc = dmalloc(h, &h->tw_x32, sizeof(float) * 2 * (size_t)h->nx))) ||
        (!h->tw_y32 && (rc = dmalloc(h, &h->tw_y32, sizeof(float) * 2 * (size_t)h->ny))) ||
        (!h->P64 && (rc = dmalloc(h, &h->P64, sizeof(double) * N))) ||
        (!h->r32 && (rc = dmalloc(h, &h->r32, sizeof(float) * N))) ||
        (!h->gauge && (rc = dmalloc(h, &h->gauge, sizeof(double) * (size_t)(pll_gauge_parts(&h->d) + 1)))) ||
        (!h->low_rx && (rc = dmalloc(h, &h->low_rx, 32 * (size_t)h->ny * h->nz))) ||
        (!h->low_y && (rc = dmalloc(h, &h->low_y, 96 * (size_t)h->nz))) ||
        (!h->low_x && (rc = dmalloc(h, &h->low_x, 96 * (size_t)h->nz))) ||
        (rc = dmalloc(h, &h->W32, sizeof(float) * N))) {
        free(tx); free(ty);
        return rc;
    }
    hipError_t e1 = hipMemcpy(h->tw_x32, tx, sizeof(float) * 2 * (size_t)h->nx, hipMemcpyHostToDevice);
    hipError_t e2 = hipMemcpy(h->tw_y32, ty, sizeof(float) * 2 * (size_t)h->ny, hipMemcpyHostToDevice);
    free(tx);
    free(ty);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        hipFree(h->W32);  /* W32 marks the set complete */
        h->W32 = NULL;
        return pll_set_error(PLL_EHIP, "copy of FP32 twiddles failed");
    }
    h->d.tw_x32 = h->tw_x32;
    h->d.tw_y32 = h->tw_y32;
    h->d.low_rx = h->low_rx;
    h->d.low_y = h->low_y;
    h->d.low_x = h->low_x;
    return PLL_OK;
}

/* The lowest modes of every FP32 solve in FP64 (pll_launch_low_solve): with
 * the residual formed in the x transform (the projection needs its FP64
 * rows), a row of the x transform inside one wavefront (its lane count a
 * power of two <= 64) and ny >= 4; PLL_REFINE_LOW=0 turns it off. */
static int low_modes_ok(const pll_ppe *h, int fused)
{
    const char *pl = getenv("PLL_REFINE_LOW");
    if (pl && atoi(pl) == 0) return 0;
    const int T = h->d.px.T;
    return fused && T >= 1 && T <= 64 && (T & (T - 1)) == 0 && h->ny >= 4 && h->L >= 2 &&
           21 * 8 * (size_t)h->nz <= 163840;
}

static int spectral32(pll_ppe *h, hipStream_t s)
{
    int rc;
    if (h->d.low_on && (rc = launch_rc(pll_launch_low_solve(&h->d, s), "lowest modes (fp64)")))
        return rc;
    if ((rc = launch_rc(pll_launch_yfwd32(&h->d, h->W32, s), "y forward (fp32)"))) return rc;
    if (h->d.low_on && (rc = launch_rc(pll_launch_low_zero(&h->d, h->W32, s), "lowest modes")))
        return rc;
    if ((rc = launch_rc(pll_launch_trid32(&h->d, h->W32, s), "z solve (fp32 storage)"))) return rc;
    return launch_rc(pll_launch_ybwd32(&h->d, h->W32, s), "y backward (fp32)");
}

static int refined_sweeps(pll_ppe *h, double *d_var, int sweeps, int fused, hipStream_t s);

int pll_ppe_solve_refined(pll_ppe *h, double *d_var, int sweeps, void *stream)
{
    if (!h) return pll_set_error(PLL_ESTATE, "solver not initialised");
    if (!d_var) return pll_set_error(PLL_EINVAL, "var is NULL");
    if (sweeps < 0 || sweeps > 64) return pll_set_error(PLL_EINVAL, "sweeps must be in [0, 64]");
    int rc;
    if ((rc = ensure_fp32(h))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const char *pf = getenv("PLL_REFINE_FUSED");
    const int fused = !(pf && atoi(pf) == 0);
    h->d.low_on = low_modes_ok(h, fused);
    rc = refined_sweeps(h, d_var, sweeps, fused, s);
    h->d.low_on = 0;
    return rc;
}

static int refined_sweeps(pll_ppe *h, double *d_var, int sweeps, int fused, hipStream_t s)
{
    int rc;
    if ((rc = launch_rc(pll_launch_xfwd32(&h->d, PLL_SRC_HALO64, d_var, h->W32, s), "x forward (fp32)")))
        return rc;
    if ((rc = spectral32(h, s))) return rc;
    if (sweeps == 0)
        return launch_rc(pll_launch_xbwd32(&h->d, PLL_DST_HALO_SET, h->W32, d_var, NULL, NULL, s),
                         "x backward (fp32)");
    if ((rc = launch_rc(pll_launch_xbwd32(&h->d, PLL_DST_DENSE_SET, h->W32, NULL, (double *)h->P64,
                                          NULL, s),
                        "x backward (fp32)")))
        return rc;
    /* Gauge: corrections keep the constant of the (0,0) mode that the first
     * solve left, i.e. the FP32 roundoff of its top-plane mean; the last
     * sweep measures the top-plane mean of P and the final write removes it. */
    double *gauge = h->d.pin_gauge ? (double *)h->gauge : NULL;
    const double *eps = gauge ? gauge + pll_gauge_parts(&h->d) : NULL;
    for (int it = 1; it <= sweeps; it++) {
        const int last = it == sweeps;
        if (fused) {
            /* residual formed in the x transform's load */
            if (last && gauge &&
                (rc = launch_rc(pll_launch_gauge(&h->d, (const double *)h->P64, 0, gauge, s), "gauge")))
                return rc;
            if ((rc = launch_rc(pll_launch_xfwd32_resid(&h->d, d_var, 1, (const double *)h->P64, 0,
                                                        h->W32, s),
                                "x forward of the residual (fp32)")))
                return rc;
        } else {
            if ((rc = launch_rc(pll_launch_resid(&h->d, d_var, (const double *)h->P64,
                                                 (float *)h->r32, last ? gauge : NULL, s),
                                "residual")))
                return rc;
            if ((rc = launch_rc(pll_launch_xfwd32(&h->d, PLL_SRC_DENSE32, h->r32, h->W32, s),
                                "x forward (fp32)")))
                return rc;
        }
        if ((rc = spectral32(h, s))) return rc;
        if ((rc = launch_rc(pll_launch_xbwd32(&h->d, last ? PLL_DST_HALO_FROM_DENSE : PLL_DST_DENSE_ADD,
                                              h->W32, d_var, (double *)h->P64, last ? eps : NULL, s),
                            "x backward (fp32)")))
            return rc;
    }
    return PLL_OK;
}

int pll_ppe_set_refinement(pll_ppe *h, int sweeps)
{
    if (!h) return pll_set_error(PLL_EINVAL, "NULL handle");
    if (sweeps < -1 || sweeps > 64) return pll_set_error(PLL_EINVAL, "sweeps must be in [-1, 64]");
    h->sweeps = sweeps;
    h->tol = 0.0;
    return sweeps >= 0 ? ensure_fp32(h) : PLL_OK;
}

/* Refinement to a tolerance.  The solution is accumulated in place in the
 * caller's haloed array and the right-hand side is kept in a dense FP64 copy
 * (the P64 buffer), so stopping after any sweep needs no extra pass.
 *
 * Stopping rule, with c_k the correction of sweep k (c_0 = the first FP32
 * solve) and |.| the max-norm: after sweep k the error of p_k is estimated as
 *   k = 1:  |c_1| / |p_1|                        (the error of p_0 bounds it)
 *   k > 1:  min(1, 10 rho_k) |c_k| / |p_k|,  rho_k = |c_k| / |c_{k-1}|.
 * rho_1 = |c_1| / |c_0| is not used: it measures the first solve's error
 * over all modes, while the error left after a sweep sits in the worst
 * conditioned modes and decays more slowly (at 768x384x256, gamma 2: ratio
 * of the first two corrections 5e-7, error after one sweep 3e-11); from the
 * second sweep on the corrections carry those modes and rho_k tracks the
 * contraction.  Each sweep reads the two norms back: one stream
 * synchronisation per sweep (none for the first FP32 solve). */
static int ensure_norm(pll_ppe *h)
{
    if (h->nrm) return PLL_OK;
    return dmalloc(h, &h->nrm, sizeof(double) * (2 + 2 * (size_t)h->ny * h->nz));
}

static int read_norms(pll_ppe *h, hipStream_t s, double out[2])
{
    if (hipMemcpyAsync(out, h->nrm, 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return pll_set_error(PLL_EHIP, "reading the refinement norms failed: %s",
                             hipGetErrorString(hipGetLastError()));
    if (!(out[0] <= DBL_MAX) || !(out[1] <= DBL_MAX))
        return pll_set_error(PLL_EINVAL, "non-finite value in the refined solve");
    return PLL_OK;
}

static int refined_tol(pll_ppe *h, double *d_var, double tol, int max_sweeps, int fused,
                       hipStream_t s);

int pll_ppe_solve_refined_tol(pll_ppe *h, double *d_var, double tol, int max_sweeps,
                              int *sweeps_done, double *err_est, void *stream)
{
    if (!h) return pll_set_error(PLL_ESTATE, "solver not initialised");
    if (!d_var) return pll_set_error(PLL_EINVAL, "var is NULL");
    if (!(tol > 0.0) || max_sweeps < 1 || max_sweeps > 64)
        return pll_set_error(PLL_EINVAL, "need tol > 0 and max_sweeps in [1, 64]");
    int rc;
    if ((rc = ensure_fp32(h)) || (rc = ensure_norm(h))) return rc;
    const char *pf = getenv("PLL_REFINE_FUSED");
    const int fused = !(pf && atoi(pf) == 0);
    h->d.low_on = low_modes_ok(h, fused);
    rc = refined_tol(h, d_var, tol, max_sweeps, fused, (hipStream_t)stream);
    h->d.low_on = 0;
    if (rc) return rc;
    if (sweeps_done) *sweeps_done = h->last_sweeps;
    if (err_est) *err_est = h->last_est;
    return PLL_OK;
}

static int refined_tol(pll_ppe *h, double *d_var, double tol, int max_sweeps, int fused,
                       hipStream_t s)
{
    int rc;
    double *F = (double *)h->P64, *nrm = (double *)h->nrm;
    if ((rc = launch_rc(pll_launch_xfwd32_copy(&h->d, d_var, F, h->W32, s), "x forward (fp32)")))
        return rc;
    if ((rc = spectral32(h, s))) return rc;
    /* the first solve's norms enter no estimate: no reduction, no read-back */
    if ((rc = launch_rc(pll_launch_xbwd32(&h->d, PLL_DST_HALO_SET, h->W32, d_var, NULL, NULL, s),
                        "x backward (fp32)")))
        return rc;
    double n[2];
    double cprev = 0.0, est = 1.0;
    double *gauge = h->d.pin_gauge ? (double *)h->gauge : NULL;
    const double *eps = gauge ? gauge + pll_gauge_parts(&h->d) : NULL;
    int it;
    for (it = 1; it <= max_sweeps; it++) {
        if (fused) {
            if (gauge && (rc = launch_rc(pll_launch_gauge(&h->d, d_var, 1, gauge, s), "gauge")))
                return rc;
            if ((rc = launch_rc(pll_launch_xfwd32_resid(&h->d, F, 0, d_var, 1, h->W32, s),
                                "x forward of the residual (fp32)")))
                return rc;
        } else {
            if ((rc = launch_rc(pll_launch_resid_hp(&h->d, F, d_var, (float *)h->r32, gauge, s),
                                "residual")))
                return rc;
            if ((rc = launch_rc(pll_launch_xfwd32(&h->d, PLL_SRC_DENSE32, h->r32, h->W32, s),
                                "x forward (fp32)")))
                return rc;
        }
        if ((rc = spectral32(h, s))) return rc;
        if ((rc = launch_rc(pll_launch_xbwd32_norm(&h->d, h->W32, d_var, eps, nrm, s),
                            "x backward (fp32)")))
            return rc;
        if ((rc = read_norms(h, s, n))) return rc;
        const double c = n[0], pmax = n[1] > 0.0 ? n[1] : 1.0;
        double f = 1.0;
        if (it > 1 && cprev > 0.0) f = 10.0 * c / cprev < 1.0 ? 10.0 * c / cprev : 1.0;
        est = f * c / pmax;
        cprev = c;
        if (est <= tol) break;
    }
    h->last_sweeps = it <= max_sweeps ? it : max_sweeps;
    h->last_est = est;
    return PLL_OK;
}

int pll_ppe_set_refinement_tol(pll_ppe *h, double tol, int max_sweeps)
{
    if (!h) return pll_set_error(PLL_EINVAL, "NULL handle");
    if (!(tol > 0.0) || max_sweeps < 1 || max_sweeps > 64)
        return pll_set_error(PLL_EINVAL, "need tol > 0 and max_sweeps in [1, 64]");
    int rc;
    if ((rc = ensure_fp32(h)) || (rc = ensure_norm(h))) return rc;
    h->sweeps = max_sweeps;
    h->tol = tol;
    return PLL_OK;
}

int pll_ppe_last_refinement(const pll_ppe *h, int *sweeps, double *err_est)
{
    if (!h) return pll_set_error(PLL_EINVAL, "NULL handle");
    if (sweeps) *sweeps = h->last_sweeps;
    if (err_est) *err_est = h->last_est;
    return PLL_OK;
}

/* ------------------------------------------------------------------ drop-in */
static pll_ppe *g_ppe = NULL;

int init_poisson_solver_status(int nx_global, int ny_global, int nz_global, double dx,
                               double dy, const double *dzf_global, const char *bctype_x,
                               const char *bctype_y, const char *bctype_z,
                               const int neighbor_xyz[3][6])
{
    if (g_ppe) return pll_set_error(PLL_ESTATE, "init_poisson_solver called twice without free");
    if (neighbor_xyz && (neighbor_xyz[2][4] >= 0 || neighbor_xyz[2][5] >= 0))
        return pll_set_error(PLL_EUNSUP,
                             "z-pencil neighbours (%d, %d) are ranks: the GPU solver holds the "
                             "whole domain and needs walls (MPI_PROC_NULL) at both z ends",
                             neighbor_xyz[2][4], neighbor_xyz[2][5]);
    return pll_ppe_create(&g_ppe, nx_global, ny_global, nz_global, dx, dy, dzf_global, bctype_x,
                          bctype_y, bctype_z);
}

static int is_device_ptr(const void *p)
{
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

int execute_poisson_solver_dev(double *d_var, void *stream)
{
    if (!g_ppe) return pll_set_error(PLL_ESTATE, "execute_poisson_solver before init");
    return pll_ppe_solve(g_ppe, d_var, stream);
}

/* Host-pointer solves (the reference's own call pattern, src/main.f90:177):
 * the interior planes of the haloed array (k = 1..nz, each a contiguous
 * (nx+2)(ny+2) block) go to the device in chunks through a ring of pinned
 * buffers: host threads copy chunk c into a pinned slot while the DMA engine
 * moves chunk c-1, and the x forward transform of a chunk starts as soon as
 * its planes have arrived.  After the y transforms and the z solve, the x
 * backward transform runs per chunk and each chunk goes back through the
 * ring the same way.  Ghost planes are neither sent nor written; the halo
 * cells inside the interior planes travel both ways unchanged.  With the
 * refined form switched on, the whole array is sent, solved and returned
 * through the same ring without the per-chunk transforms.
 * PLL_HOST_PIPE=0 selects one plain hipMemcpy each way instead. */
static int host_setup(pll_ppe *h)
{
    if (h->host_ready) return PLL_OK;
    const size_t plane = sizeof(double) * (size_t)(h->nx + 2) * (h->ny + 2);
    /* slots of at most 20 MiB (whole planes): small chunks start the first
     * transform and the last download sooner (768x384x256: 8 planes 24.0 ms
     * per solve, 26 planes 25.3 ms) */
    size_t cp = ((size_t)20 << 20) / plane;
    if (cp < 1) cp = 1;
    if (cp > (size_t)h->nz) cp = (size_t)h->nz;
    const char *pc = getenv("PLL_HOST_CHUNK"); /* tuning: planes per chunk */
    if (pc && atoi(pc) > 0) cp = (size_t)atoi(pc) < (size_t)h->nz ? (size_t)atoi(pc) : (size_t)h->nz;
    for (int i = 0; i < PLL_HOST_SLOTS; i++)
        if (!h->pin[i]) HIPCHK(hipHostMalloc(&h->pin[i], cp * plane, hipHostMallocDefault));
    h->slot_planes = cp;
    HIPCHK(hipStreamCreateWithFlags(&h->hs_copy, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&h->hs_comp, hipStreamNonBlocking));
    for (int i = 0; i < PLL_HOST_SLOTS; i++) {
        HIPCHK(hipEventCreateWithFlags(&h->ev_up[i], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&h->ev_x[i], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&h->ev_dn[i], hipEventDisableTiming));
    }
    int nt = 8;
    const char *pt = getenv("PLL_HOST_THREADS"); /* host copy threads */
    if (pt && atoi(pt) > 0) nt = atoi(pt);
    if (!h->pool && pll_copy_pool_create(&h->pool, nt) != 0)
        return pll_set_error(PLL_ENOMEM, "host copy threads could not be created");
    h->host_ready = 1;
    return PLL_OK;
}

static int host_solve(pll_ppe *h, double *var)
{
    int rc;
    if ((rc = host_setup(h))) return rc;
    const long long off1 = (long long)(h->nx + 2) * (h->ny + 2);
    const size_t plane = sizeof(double) * (size_t)off1;
    const int nz = h->nz, ny = h->ny, cp = (int)h->slot_planes;
    const int nch = (nz + cp - 1) / cp;
    const int fp64 = h->sweeps < 0 && !(h->tol > 0.0);
    hipStream_t cs = h->hs_copy, ks = h->hs_comp;
    /* upload (and x forward per chunk) */
    for (int c = 0; c < nch; c++) {
        const int k0 = c * cp, k1 = k0 + cp < nz ? k0 + cp : nz, sl = c % PLL_HOST_SLOTS;
        const size_t bytes = plane * (size_t)(k1 - k0);
        if (c >= PLL_HOST_SLOTS) HIPCHK(hipEventSynchronize(h->ev_up[sl]));
        pll_copy_pool_memcpy(h->pool, h->pin[sl], var + (k0 + 1) * off1, bytes);
        HIPCHK(hipMemcpyAsync(h->stage + (k0 + 1) * off1, h->pin[sl], bytes,
                              hipMemcpyHostToDevice, cs));
        HIPCHK(hipEventRecord(h->ev_up[sl], cs));
        HIPCHK(hipStreamWaitEvent(ks, h->ev_up[sl], 0));
        if (fp64 && (rc = launch_rc(pll_launch_xfwd_rows(&h->d, h->stage, h->W, (long long)k0 * ny,
                                                         (long long)k1 * ny, ks),
                                    "x forward")))
            return rc;
    }
    if (fp64) {
        if ((rc = launch_rc(pll_launch_yfwd(&h->d, h->W, ks), "y forward")) ||
            (rc = launch_rc(pll_launch_trid(&h->d, h->W, ks), "z solve")) ||
            (rc = launch_rc(pll_launch_ybwd(&h->d, h->W, ks), "y backward")))
            return rc;
    } else if ((rc = pll_ppe_solve(h, h->stage, ks))) {
        return rc;
    }
    /* x backward per chunk (FP64) and download; a slot is reused only after
     * the host has copied its previous chunk out */
    for (int c = 0; c < nch + PLL_HOST_SLOTS; c++) {
        const int cd = c - PLL_HOST_SLOTS; /* chunk to drain before slot reuse */
        if (cd >= 0) {
            const int k0 = cd * cp, k1 = k0 + cp < nz ? k0 + cp : nz, sl = cd % PLL_HOST_SLOTS;
            HIPCHK(hipEventSynchronize(h->ev_dn[sl]));
            pll_copy_pool_memcpy(h->pool, var + (k0 + 1) * off1, h->pin[sl],
                                 plane * (size_t)(k1 - k0));
        }
        if (c >= nch) continue;
        const int k0 = c * cp, k1 = k0 + cp < nz ? k0 + cp : nz, sl = c % PLL_HOST_SLOTS;
        if (fp64 && (rc = launch_rc(pll_launch_xbwd_rows(&h->d, h->W, h->stage, (long long)k0 * ny,
                                                         (long long)k1 * ny, ks),
                                    "x backward")))
            return rc;
        HIPCHK(hipEventRecord(h->ev_x[sl], ks));
        HIPCHK(hipStreamWaitEvent(cs, h->ev_x[sl], 0));
        HIPCHK(hipMemcpyAsync(h->pin[sl], h->stage + (k0 + 1) * off1, plane * (size_t)(k1 - k0),
                              hipMemcpyDeviceToHost, cs));
        HIPCHK(hipEventRecord(h->ev_dn[sl], cs));
    }
    return PLL_OK;
}

int execute_poisson_solver_status(double *var)
{
    pll_ppe *h = g_ppe;
    if (!h) return pll_set_error(PLL_ESTATE, "execute_poisson_solver before init");
    if (!var) return pll_set_error(PLL_EINVAL, "var is NULL");
    int rc;
    if (is_device_ptr(var)) {
        if ((rc = pll_ppe_solve(h, var, NULL))) return rc;
        HIPCHK(hipStreamSynchronize(NULL));
        return PLL_OK;
    }
    const size_t bytes = sizeof(double) * (size_t)(h->nx + 2) * (h->ny + 2) * (h->nz + 2);
    if (!h->stage) {
        if ((rc = dmalloc(h, (void **)&h->stage, bytes))) return rc;
        h->stage_bytes = bytes;
    }
    const char *pp = getenv("PLL_HOST_PIPE");
    if (!(pp && atoi(pp) == 0)) return host_solve(h, var);
    HIPCHK(hipMemcpy(h->stage, var, bytes, hipMemcpyHostToDevice));
    if ((rc = pll_ppe_solve(h, h->stage, NULL))) return rc;
    HIPCHK(hipMemcpy(var, h->stage, bytes, hipMemcpyDeviceToHost));
    return PLL_OK;
}

int set_poisson_solver_refinement(int sweeps)
{
    if (!g_ppe) return pll_set_error(PLL_ESTATE, "set_poisson_solver_refinement before init");
    return pll_ppe_set_refinement(g_ppe, sweeps);
}

int free_poisson_solver_status(void)
{
    int rc = pll_ppe_destroy(g_ppe);
    g_ppe = NULL;
    return rc;
}

static void abort_on(int rc, const char *what)
{
    if (rc == PLL_OK) return;
    fprintf(stderr, "PowerLLEL-MI355X.ERROR: %s: %s (code %d)\n", what, pll_last_error(), rc);
    fflush(stderr);
    abort();
}

void init_poisson_solver(int nx_global, int ny_global, int nz_global, double dx, double dy,
                         double *dzf_global, char bctype_x[], char bctype_y[], char bctype_z[],
                         int neighbor_xyz[3][6])
{
    abort_on(init_poisson_solver_status(nx_global, ny_global, nz_global, dx, dy, dzf_global,
                                        bctype_x, bctype_y, bctype_z,
                                        (const int(*)[6])neighbor_xyz),
             "init_poisson_solver");
}

void execute_poisson_solver(double *var)
{
    abort_on(execute_poisson_solver_status(var), "execute_poisson_solver");
}

void free_poisson_solver(void) { abort_on(free_poisson_solver_status(), "free_poisson_solver"); }

/* Host-only query of the FFT plan chosen for length L (no device needed):
 * points per thread, stage count and radices (rad: PLL_MAX_STAGES ints). */
int pll_fft_plan_info(int L, int *E, int *nst, int *rad)
{
    pll_fftplan p;
    const int rc = make_plan(L, &p);
    if (rc) return rc;
    *E = p.E;
    *nst = p.nst;
    for (int s = 0; s < p.nst; s++) rad[s] = p.rad[s];
    return PLL_OK;
}
