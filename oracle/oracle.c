/*
 * Oracle: CPU restatement of PowerLLEL_Channel's time-step hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h for the rules and the reference
 * file:line map).  Built with -ffp-contract=off so every product and sum is
 * rounded separately, in the order the reference's source writes it.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#ifdef _OPENMP
#define NT num_threads(orc_get_threads())
#endif

static double now_s(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

static void die(const char *msg)
{
    fprintf(stderr, "oracle: %s\n", msg);
    abort();
}

/* ===================================================================== */
/* Complex FFT: Stockham autosort, mixed radix 4/2/3/5 + generic odd      */
/* prime.  Twiddles are exp(-2 pi i t / n) computed in long double.       */
/* ===================================================================== */
typedef struct {
    int n;
    int nf;
    int fac[64];
    double *twr, *twi; /* [n] */
} cplan_t;

static void cplan_init(cplan_t *p, int n)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[p->nf++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[p->nf++] = 5; m /= 5; }
    for (int f = 7; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
    }
    p->twr = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    p->twi = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; t++) {
        long double ang = twopi * (long double)t / (long double)n;
        p->twr[t] = (double)cosl(ang);
        p->twi[t] = (double)(-sinl(ang));
    }
}

static void cplan_free(cplan_t *p)
{
    free(p->twr);
    free(p->twi);
    p->twr = p->twi = NULL;
}

/* sign = -1 forward, +1 backward (unnormalised).  (re,im) in/out; (wr,wi)
 * scratch of length n. */
static void cfft_exec(const cplan_t *p, double *re, double *im, double *wr,
                      double *wi, int sign)
{
    const int n = p->n;
    if (n <= 1) return;
    double *ir = re, *ii = im, *orr = wr, *oi = wi;
    int Ns = 1;
    const double s3 = 0.86602540378443864676372317075294; /* sqrt(3)/2 */
    for (int s = 0; s < p->nf; s++) {
        const int R = p->fac[s];
        const int m = n / R;
        const int step = n / (Ns * R);
        double xr[64], xi[64], yr[64], yi[64];
        if (R > 64) die("radix too large");
        for (int j = 0; j < m; j++) {
            const int k = j % Ns;
            for (int r = 0; r < R; r++) {
                double ar = ir[j + r * m], ai = ii[j + r * m];
                if (r > 0 && k > 0) {
                    int t = r * k * step;
                    double cr = p->twr[t], ci = sign < 0 ? p->twi[t] : -p->twi[t];
                    double tr = ar * cr - ai * ci;
                    double ti = ar * ci + ai * cr;
                    ar = tr;
                    ai = ti;
                }
                xr[r] = ar;
                xi[r] = ai;
            }
            if (R == 2) {
                yr[0] = xr[0] + xr[1]; yi[0] = xi[0] + xi[1];
                yr[1] = xr[0] - xr[1]; yi[1] = xi[0] - xi[1];
            } else if (R == 4) {
                double t0r = xr[0] + xr[2], t0i = xi[0] + xi[2];
                double t1r = xr[0] - xr[2], t1i = xi[0] - xi[2];
                double t2r = xr[1] + xr[3], t2i = xi[1] + xi[3];
                double t3r = xr[1] - xr[3], t3i = xi[1] - xi[3];
                /* sign*i*(t3) = (-sign*t3i, sign*t3r) */
                yr[0] = t0r + t2r; yi[0] = t0i + t2i;
                yr[2] = t0r - t2r; yi[2] = t0i - t2i;
                yr[1] = t1r - sign * t3i; yi[1] = t1i + sign * t3r;
                yr[3] = t1r + sign * t3i; yi[3] = t1i - sign * t3r;
            } else if (R == 3) {
                double sr = xr[1] + xr[2], si = xi[1] + xi[2];
                double dr = xr[1] - xr[2], di = xi[1] - xi[2];
                double mr = xr[0] - 0.5 * sr, mi = xi[0] - 0.5 * si;
                /* sign*i*s3*(d) = (-sign*s3*di, sign*s3*dr) */
                yr[0] = xr[0] + sr; yi[0] = xi[0] + si;
                yr[1] = mr - sign * s3 * di; yi[1] = mi + sign * s3 * dr;
                yr[2] = mr + sign * s3 * di; yi[2] = mi - sign * s3 * dr;
            } else {
                /* generic DFT of length R using the n-point table */
                const int tstep = n / R;
                for (int q = 0; q < R; q++) {
                    double accr = 0.0, acci = 0.0;
                    for (int r = 0; r < R; r++) {
                        int t = ((r * q) % R) * tstep;
                        double cr = p->twr[t], ci = sign < 0 ? p->twi[t] : -p->twi[t];
                        accr += xr[r] * cr - xi[r] * ci;
                        acci += xr[r] * ci + xi[r] * cr;
                    }
                    yr[q] = accr;
                    yi[q] = acci;
                }
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) {
                orr[base + r * Ns] = yr[r];
                oi[base + r * Ns] = yi[r];
            }
        }
        double *t;
        t = ir; ir = orr; orr = t;
        t = ii; ii = oi; oi = t;
        Ns *= R;
    }
    if (ir != re) {
        memcpy(re, ir, sizeof(double) * (size_t)n);
        memcpy(im, ii, sizeof(double) * (size_t)n);
    }
}

/* ===================================================================== */
/* Real transforms in FFTW's halfcomplex order:                          */
/*   out = (Re X_0, Re X_1, ..., Re X_{n/2}, Im X_{(n+1)/2-1}, ..., Im X_1) */
/* with X_m = sum_j x_j exp(-2 pi i j m / n).  HC2R is the unnormalised   */
/* inverse (HC2R(R2HC(x)) = n x).  Even n use the half-length complex     */
/* FFT of z_m = x_{2m} + i x_{2m+1}; odd n a full-length complex FFT.     */
/* ===================================================================== */
typedef struct {
    int n, L;
    cplan_t cp;        /* length L (even n) or n (odd n) */
    double *er, *ei;   /* exp(-2 pi i m / n), m in [0, n/2] */
} rplan_t;

static void rplan_init(rplan_t *p, int n)
{
    p->n = n;
    p->L = (n % 2 == 0) ? n / 2 : n;
    cplan_init(&p->cp, p->L);
    int ne = n / 2 + 1;
    p->er = (double *)malloc(sizeof(double) * (size_t)ne);
    p->ei = (double *)malloc(sizeof(double) * (size_t)ne);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < ne; m++) {
        long double ang = twopi * (long double)m / (long double)n;
        p->er[m] = (double)cosl(ang);
        p->ei[m] = (double)(-sinl(ang));
    }
}

static void rplan_free(rplan_t *p)
{
    cplan_free(&p->cp);
    free(p->er);
    free(p->ei);
}

/* work: 4*L doubles */
static void r2hc_exec(const rplan_t *p, const double *in, double *out, double *work)
{
    const int n = p->n, L = p->L;
    double *zr = work, *zi = work + L, *wr = work + 2 * L, *wi = work + 3 * L;
    if (n == 1) { out[0] = in[0]; return; }
    if (n % 2 == 0) {
        for (int m = 0; m < L; m++) { zr[m] = in[2 * m]; zi[m] = in[2 * m + 1]; }
        cfft_exec(&p->cp, zr, zi, wr, wi, -1);
        for (int m = 0; m <= L; m++) {
            int m1 = m % L, m2 = (L - m) % L;
            double Zr = zr[m1], Zi = zi[m1];
            double Cr = zr[m2], Ci = -zi[m2];          /* conj Z_{L-m} */
            double Er = 0.5 * (Zr + Cr), Ei = 0.5 * (Zi + Ci);
            double Dr = Zr - Cr, Di = Zi - Ci;          /* Z - conj   */
            double Or = 0.5 * Di, Oi = -0.5 * Dr;       /* -i/2 * D   */
            double wre = p->er[m], wim = p->ei[m];
            double Xr = Er + (wre * Or - wim * Oi);
            double Xi = Ei + (wre * Oi + wim * Or);
            out[m] = Xr;
            if (m > 0 && m < L) out[n - m] = Xi;
        }
    } else {
        for (int m = 0; m < n; m++) { zr[m] = in[m]; zi[m] = 0.0; }
        cfft_exec(&p->cp, zr, zi, wr, wi, -1);
        for (int m = 0; m <= n / 2; m++) out[m] = zr[m];
        for (int m = 1; m <= n / 2; m++) out[n - m] = zi[m];
    }
}

static void hc2r_exec(const rplan_t *p, const double *in, double *out, double *work)
{
    const int n = p->n, L = p->L;
    double *zr = work, *zi = work + L, *wr = work + 2 * L, *wi = work + 3 * L;
    if (n == 1) { out[0] = in[0]; return; }
    if (n % 2 == 0) {
        for (int k = 0; k < L; k++) {
            double Xr = in[k], Xi = (k == 0) ? 0.0 : in[n - k];
            int q = L - k; /* 1..L */
            double Cr = in[q], Ci = (q == L) ? 0.0 : -in[n - q]; /* conj X_{L-k} */
            double Sr = Xr + Cr, Si = Xi + Ci;
            double Dr = Xr - Cr, Di = Xi - Ci;
            /* e^{+2 pi i k/n} = conj(er,ei) */
            double cr = p->er[k], ci = -p->ei[k];
            double Tr = cr * Dr - ci * Di, Ti = cr * Di + ci * Dr;
            /* Z' = S + i T */
            zr[k] = Sr - Ti;
            zi[k] = Si + Tr;
        }
        cfft_exec(&p->cp, zr, zi, wr, wi, +1);
        for (int m = 0; m < L; m++) { out[2 * m] = zr[m]; out[2 * m + 1] = zi[m]; }
    } else {
        zr[0] = in[0]; zi[0] = 0.0;
        for (int m = 1; m <= n / 2; m++) {
            zr[m] = in[m]; zi[m] = in[n - m];
            zr[n - m] = in[m]; zi[n - m] = -in[n - m];
        }
        cfft_exec(&p->cp, zr, zi, wr, wi, +1);
        for (int m = 0; m < n; m++) out[m] = zr[m];
    }
}

void orc_r2hc(int n, const double *in, double *out)
{
    rplan_t p;
    rplan_init(&p, n);
    double *work = (double *)malloc(sizeof(double) * 4 * (size_t)p.L);
    r2hc_exec(&p, in, out, work);
    free(work);
    rplan_free(&p);
}

void orc_hc2r(int n, const double *in, double *out)
{
    rplan_t p;
    rplan_init(&p, n);
    double *work = (double *)malloc(sizeof(double) * 4 * (size_t)p.L);
    hc2r_exec(&p, in, out, work);
    free(work);
    rplan_free(&p);
}

/* ===================================================================== */
/* Mesh: src/mod_mesh.f90:120-162 (initMeshByFunc, tanh) and :64-72.       */
/* ===================================================================== */
void orc_mesh_tanh(int nz, double lz, double stretch_ratio, double *dzf_global)
{
    if (stretch_ratio == 0.0) { /* mod_mesh.f90:130-132 */
        for (int k = 0; k <= nz + 1; k++) dzf_global[k] = lz / nz;
        return;
    }
    double *zf = (double *)malloc(sizeof(double) * (size_t)(nz + 1));
    zf[0] = 0.0;
    for (int k = 1; k <= nz; k++) { /* mod_mesh.f90:143-146 */
        double z0 = k / (1.0 * nz);
        zf[k] = 0.5 * (1.0 + tanh((z0 - 0.5) * stretch_ratio) / tanh(stretch_ratio / 2.0));
    }
    for (int k = 1; k <= nz; k++) dzf_global[k] = lz * (zf[k] - zf[k - 1]);
    dzf_global[0] = dzf_global[1];
    dzf_global[nz + 1] = dzf_global[nz];
    free(zf);
}

void orc_mesh_derived(int nz, double lz, const double *dzf, double *dzf_inv,
                      double *dzc, double *dzc_inv, double *dzflzi)
{
    for (int k = 0; k <= nz + 1; k++) dzf_inv[k] = 1.0 / dzf[k]; /* mod_mesh.f90:64-67 */
    for (int k = 0; k <= nz; k++) {                                /* :68-71 */
        dzc[k] = 0.5 * (dzf[k] + dzf[k + 1]);
        dzc_inv[k] = 1.0 / dzc[k];
    }
    dzc[nz + 1] = 0.0;
    dzc_inv[nz + 1] = 0.0;
    for (int k = 0; k <= nz + 1; k++) dzflzi[k] = dzf[k] / lz; /* :72 */
}

/* ===================================================================== */
/* PPE pieces: mod_fft_fftw.inc:175-198 and mod_poisson_solver.c:40-103   */
/* ===================================================================== */
void orc_eigen_values(int ist, int isz, int isz_global, const char *bctype,
                      double *lambda)
{
    const double pi = acos(-1.0);
    int ien = ist + isz - 1;
    if (strcmp(bctype, "PP") == 0) {
        for (int i = ist; i <= ien; i++)
            lambda[i - ist] = 2.0 * (cos(2.0 * pi * (i - 1.0) / isz_global) - 1.0);
    } else if (strcmp(bctype, "NN") == 0) {
        for (int i = ist; i <= ien; i++)
            lambda[i - ist] = 2.0 * (cos(pi * (i - 1.0) / isz_global) - 1.0);
    } else if (strcmp(bctype, "DD") == 0) {
        for (int i = ist; i <= ien; i++)
            lambda[i - ist] = 2.0 * (cos(pi * (i - 0.0) / isz_global) - 1.0);
    } else if (strcmp(bctype, "ND") == 0 || strcmp(bctype, "DN") == 0) {
        for (int i = ist; i <= ien; i++)
            lambda[i - ist] = 2.0 * (cos(pi * (2 * i - 1.0) / (2.0 * isz_global)) - 1.0);
    } else {
        die("unknown bctype in orc_eigen_values");
    }
}

/* dzf points at dzf(0) of the local z range: n+2 entries.  A neighbour value
 * < 0 means MPI_PROC_NULL (a wall). */
void orc_set_trid_coeff(int n, const double *dzf, const char *bctype, char c_or_f,
                        const int neighbor[2], double *a, double *b, double *c)
{
    const int dzf_st = -1;
    if (c_or_f == 'c') {
        for (int k = 0; k < n; k++) {
            a[k] = 2.0 / (dzf[k - dzf_st] * (dzf[k - 1 - dzf_st] + dzf[k - dzf_st]));
            c[k] = 2.0 / (dzf[k - dzf_st] * (dzf[k + 1 - dzf_st] + dzf[k - dzf_st]));
        }
    } else if (c_or_f == 'f') {
        for (int k = 0; k < n; k++) {
            a[k] = 2.0 / (dzf[k - dzf_st] * (dzf[k + 1 - dzf_st] + dzf[k - dzf_st]));
            c[k] = 2.0 / (dzf[k + 1 - dzf_st] * (dzf[k + 1 - dzf_st] + dzf[k - dzf_st]));
        }
    } else {
        die("c_or_f must be 'c' or 'f'");
    }
    for (int i = 0; i < n; i++) b[i] = -a[i] - c[i];

    for (int side = 0; side < 2; side++) {
        if (neighbor[side] >= 0) continue; /* interior neighbour: no correction */
        char bc = bctype[side];
        double factor;
        if (bc == 'P') factor = 0.0;
        else if (bc == 'D') factor = -1.0;
        else if (bc == 'N') factor = 1.0;
        else die("bad z boundary type");
        if (side == 0) {
            if (c_or_f == 'c' || bc == 'N') {
                b[0] += factor * a[0];
                a[0] *= fabs(factor) - 1.0;
            }
        } else {
            if (c_or_f == 'c' || bc == 'N') {
                b[n - 1] += factor * c[n - 1];
                c[n - 1] *= fabs(factor) - 1.0;
            }
        }
    }
}

/* ===================================================================== */
/* PPE solver: mod_poisson_solver.c:178-309, 311-359, 468-605             */
/* ===================================================================== */
struct orc_ppe {
    int nx, ny, nz;
    double dx, dy;
    int pin_gauge, singular;
    double normfactor;
    double *work;  /* nx*ny*nz, x fastest */
    double *a, *c; /* nz */
    double *bz;    /* nz: b_tmp before adding the eigenvalues */
    double *b;     /* nx*ny*nz factorised pivots */
    double *lxy;   /* nx*ny */
    rplan_t px, py;
    double tstage[5];
};

orc_ppe_t *orc_ppe_init(int nx, int ny, int nz, double dx, double dy,
                        const double *dzf_global, const char *bctype_x,
                        const char *bctype_y, const char *bctype_z, int pin_gauge)
{
    if (strcmp(bctype_x, "PP") != 0 || strcmp(bctype_y, "PP") != 0)
        die("oracle PPE supports periodic x and y only");
    orc_ppe_t *h = (orc_ppe_t *)calloc(1, sizeof(orc_ppe_t));
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->dx = dx; h->dy = dy;
    h->pin_gauge = pin_gauge;
    h->singular = (bctype_z[0] == 'N' && bctype_z[1] == 'N');
    const size_t N = (size_t)nx * ny * nz;
    h->work = (double *)malloc(sizeof(double) * N);
    h->b = (double *)malloc(sizeof(double) * N);
    h->a = (double *)malloc(sizeof(double) * (size_t)nz);
    h->c = (double *)malloc(sizeof(double) * (size_t)nz);
    h->bz = (double *)malloc(sizeof(double) * (size_t)nz);
    h->lxy = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    rplan_init(&h->px, nx);
    rplan_init(&h->py, ny);
    /* mod_fft_fftw.inc:154: normfactor = 1/(1*nx*1*ny) for PP,PP */
    h->normfactor = 1.0 / (1.0 * nx * 1.0 * ny);

    double *lx = (double *)malloc(sizeof(double) * (size_t)nx);
    double *ly = (double *)malloc(sizeof(double) * (size_t)ny);
    orc_eigen_values(1, nx, nx, bctype_x, lx); /* :230-237 */
    for (int i = 0; i < nx; i++) lx[i] /= (dx * dx);
    orc_eigen_values(1, ny, ny, bctype_y, ly);
    for (int i = 0; i < ny; i++) ly[i] /= (dy * dy);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) h->lxy[i + (size_t)j * nx] = lx[i] + ly[j];

    const int nb[2] = {-1, -1}; /* walls at both z ends (one rank) */
    orc_set_trid_coeff(nz, dzf_global, bctype_z, 'c', nb, h->a, h->bz, h->c);
    const size_t P = (size_t)nx * ny;
#pragma omp parallel for schedule(static) NT
    for (size_t ij = 0; ij < P; ij++) {
        h->b[ij] = h->bz[0] + h->lxy[ij];
        for (int k = 1; k < nz; k++) { /* :257-275 */
            size_t e = ij + (size_t)k * P;
            h->b[e] = h->bz[k] + h->lxy[ij];
            double a_tmp = h->a[k] / h->b[e - P];
            h->b[e] -= (a_tmp * h->c[k - 1]);
        }
    }
    free(lx);
    free(ly);
    return h;
}

void orc_ppe_free(orc_ppe_t *h)
{
    if (!h) return;
    free(h->work); free(h->b); free(h->a); free(h->c); free(h->bz); free(h->lxy);
    rplan_free(&h->px);
    rplan_free(&h->py);
    free(h);
}

void orc_ppe_stage_times(const orc_ppe_t *h, double t[5])
{
    for (int i = 0; i < 5; i++) t[i] = h->tstage[i];
}

#define YBLK 16
void orc_ppe_solve(orc_ppe_t *h, double *var)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const size_t off1 = (size_t)(nx + 2) * (ny + 2), off2 = (size_t)(nx + 2);
    const size_t P = (size_t)nx * ny;
    double *xp = h->work;
    double t0 = now_s();

    /* --Copy & Fwd X-FFT (:484-494) */
#pragma omp parallel NT
    {
        double *wk = (double *)malloc(sizeof(double) * (4 * (size_t)h->px.L + nx));
        double *row = wk + 4 * (size_t)h->px.L;
#pragma omp for schedule(static)
        for (int k = 0; k < nz; k++) {
            for (int j = 0; j < ny; j++) {
                const double *src = &var[(size_t)(k + 1) * off1 + (size_t)(j + 1) * off2 + 1];
                memcpy(row, src, sizeof(double) * (size_t)nx);
                r2hc_exec(&h->px, row, &xp[(size_t)k * P + (size_t)j * nx], wk);
            }
        }
        free(wk);
    }
    double t1 = now_s();

    /* --Forward Y-FFT (:510-514): stride nx, one column per x slot */
#pragma omp parallel NT
    {
        double *wk = (double *)malloc(sizeof(double) * (4 * (size_t)h->py.L + 2 * (size_t)ny * YBLK));
        double *col = wk + 4 * (size_t)h->py.L, *res = col + (size_t)ny * YBLK;
#pragma omp for schedule(static)
        for (int k = 0; k < nz; k++) {
            double *pl = &xp[(size_t)k * P];
            for (int i0 = 0; i0 < nx; i0 += YBLK) {
                int nb = (nx - i0 < YBLK) ? nx - i0 : YBLK;
                for (int j = 0; j < ny; j++)
                    for (int b = 0; b < nb; b++) col[(size_t)b * ny + j] = pl[(size_t)j * nx + i0 + b];
                for (int b = 0; b < nb; b++) r2hc_exec(&h->py, &col[(size_t)b * ny], &res[(size_t)b * ny], wk);
                for (int j = 0; j < ny; j++)
                    for (int b = 0; b < nb; b++) pl[(size_t)j * nx + i0 + b] = res[(size_t)b * ny + j];
            }
        }
        free(wk);
    }
    double t2 = now_s();

    /* --Solve trid: ssolve_trid (:311-359); each (i,j) column is independent,
     * so the loops are interchanged (j outer) with identical arithmetic. */
    const double *a = h->a, *c = h->c, *b = h->b;
#pragma omp parallel for schedule(static) NT
    for (int j = 0; j < ny; j++) {
        for (int k = 1; k < nz; k++) {
            double *vd = &xp[(size_t)k * P + (size_t)j * nx];
            const double *vu = &xp[(size_t)(k - 1) * P + (size_t)j * nx];
            const double *bp = &b[(size_t)(k - 1) * P + (size_t)j * nx];
            for (int i = 0; i < nx; i++) {
                double a_tmp = a[k] / bp[i];
                vd[i] -= a_tmp * vu[i];
            }
        }
        {
            const double *bp = &b[(size_t)(nz - 1) * P + (size_t)j * nx];
            double *vp = &xp[(size_t)(nz - 1) * P + (size_t)j * nx];
            for (int i = 0; i < nx; i++) {
                if (h->pin_gauge && h->singular && i == 0 && j == 0) vp[i] = 0.0;
                else if (bp[i] != 0) vp[i] /= bp[i];
                else vp[i] = 0;
            }
        }
        for (int k = nz - 2; k >= 0; k--) {
            double *vd = &xp[(size_t)k * P + (size_t)j * nx];
            const double *vu = &xp[(size_t)(k + 1) * P + (size_t)j * nx];
            const double *bp = &b[(size_t)k * P + (size_t)j * nx];
            for (int i = 0; i < nx; i++) vd[i] = (vd[i] - c[k] * vu[i]) / bp[i];
        }
    }
    double t3 = now_s();

    /* --Backward Y-FFT (:568-571) */
#pragma omp parallel NT
    {
        double *wk = (double *)malloc(sizeof(double) * (4 * (size_t)h->py.L + 2 * (size_t)ny * YBLK));
        double *col = wk + 4 * (size_t)h->py.L, *res = col + (size_t)ny * YBLK;
#pragma omp for schedule(static)
        for (int k = 0; k < nz; k++) {
            double *pl = &xp[(size_t)k * P];
            for (int i0 = 0; i0 < nx; i0 += YBLK) {
                int nb = (nx - i0 < YBLK) ? nx - i0 : YBLK;
                for (int j = 0; j < ny; j++)
                    for (int bb = 0; bb < nb; bb++) col[(size_t)bb * ny + j] = pl[(size_t)j * nx + i0 + bb];
                for (int bb = 0; bb < nb; bb++) hc2r_exec(&h->py, &col[(size_t)bb * ny], &res[(size_t)bb * ny], wk);
                for (int j = 0; j < ny; j++)
                    for (int bb = 0; bb < nb; bb++) pl[(size_t)j * nx + i0 + bb] = res[(size_t)bb * ny + j];
            }
        }
        free(wk);
    }
    double t4 = now_s();

    /* --Bwd X-FFT & Copy (:587-599) */
#pragma omp parallel NT
    {
        double *wk = (double *)malloc(sizeof(double) * (4 * (size_t)h->px.L + nx));
        double *row = wk + 4 * (size_t)h->px.L;
#pragma omp for schedule(static)
        for (int k = 0; k < nz; k++) {
            for (int j = 0; j < ny; j++) {
                hc2r_exec(&h->px, &xp[(size_t)k * P + (size_t)j * nx], row, wk);
                double *dst = &var[(size_t)(k + 1) * off1 + (size_t)(j + 1) * off2 + 1];
                for (int i = 0; i < nx; i++) dst[i] = row[i] * h->normfactor;
            }
        }
        free(wk);
    }
    double t5 = now_s();
    h->tstage[0] = t1 - t0; h->tstage[1] = t2 - t1; h->tstage[2] = t3 - t2;
    h->tstage[3] = t4 - t3; h->tstage[4] = t5 - t4;
}

void orc_ppe_apply(orc_ppe_t *h, const double *var, double *out)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const size_t off1 = (size_t)(nx + 2) * (ny + 2), off2 = (size_t)(nx + 2);
    const double idx2 = 1.0 / (h->dx * h->dx), idy2 = 1.0 / (h->dy * h->dy);
#pragma omp parallel for schedule(static) NT
    for (int k = 0; k < nz; k++) {
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++) {
                int ip = (i + 1) % nx, im = (i + nx - 1) % nx;
                int jp = (j + 1) % ny, jm = (j + ny - 1) % ny;
#define PV(ii, jj, kk) var[(size_t)((kk) + 1) * off1 + (size_t)((jj) + 1) * off2 + (ii) + 1]
                double pc = PV(i, j, k);
                double sx = (PV(ip, j, k) - 2.0 * pc + PV(im, j, k)) * idx2;
                double sy = (PV(i, jp, k) - 2.0 * pc + PV(i, jm, k)) * idy2;
                double sz = h->bz[k] * pc;
                if (k > 0) sz += h->a[k] * PV(i, j, k - 1);
                if (k < nz - 1) sz += h->c[k] * PV(i, j, k + 1);
                out[(size_t)k * nx * ny + (size_t)j * nx + i] = sx + sy + sz;
#undef PV
            }
        }
    }
}

/* ===================================================================== */
/* Step kernels                                                           */
/* ===================================================================== */
#define IX(g, i, j, k) ((size_t)(i) + (size_t)((g)->nx + 2) * ((size_t)(j) + (size_t)((g)->ny + 2) * (size_t)(k)))

void orc_transform2crf(const orc_grid_t *g, double vel_crf, double *vel)
{ /* mod_calcVel.f90:26-50 */
#pragma omp parallel for schedule(static) NT
    for (int k = 1; k <= g->nz; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) vel[IX(g, i, j, k)] = vel[IX(g, i, j, k)] - vel_crf;
}

/* The three momentum right-hand sides, written exactly as
 * mod_calcVel.f90:74-138 (RK1) and :262-326 (RK2).  q = velocity the
 * operator is applied to. */
#define U(i, j, k) qu[IX(g, i, j, k)]
#define V(i, j, k) qv[IX(g, i, j, k)]
#define W(i, j, k) qw[IX(g, i, j, k)]
static inline void mom_rhs(const orc_grid_t *g, const double *qu, const double *qv,
                           const double *qw, int i, int j, int k, double u_crf,
                           double *fu, double *fv, double *fw)
{
    const double dx_inv = g->dx_inv, dy_inv = g->dy_inv, re_inv = g->re_inv;
    const double *dzf = g->dzf, *dzf_inv = g->dzf_inv, *dzc_inv = g->dzc_inv;
    double uw, ue, us, un, ub, ut, ww, we, ws, wn, wb, wt, vw, ve, vs, vn, vb, vt;
    double dqw, dqe, dqs, dqn, dqb, dqt, conv, visc;

    ue = (U(i, j, k) + U(i + 1, j, k));
    uw = (U(i - 1, j, k) + U(i, j, k));
    un = (U(i, j, k) + U(i, j + 1, k));
    us = (U(i, j - 1, k) + U(i, j, k));
    ut = (U(i, j, k) * dzf[k + 1] + U(i, j, k + 1) * dzf[k]) * dzc_inv[k];
    ub = (U(i, j, k) * dzf[k - 1] + U(i, j, k - 1) * dzf[k]) * dzc_inv[k - 1];
    vn = (V(i, j, k) + V(i + 1, j, k));
    vs = (V(i, j - 1, k) + V(i + 1, j - 1, k));
    wt = (W(i, j, k) + W(i + 1, j, k));
    wb = (W(i, j, k - 1) + W(i + 1, j, k - 1));
    dqe = (U(i + 1, j, k) - U(i, j, k)) * dx_inv;
    dqw = (U(i, j, k) - U(i - 1, j, k)) * dx_inv;
    dqn = (U(i, j + 1, k) - U(i, j, k)) * dy_inv;
    dqs = (U(i, j, k) - U(i, j - 1, k)) * dy_inv;
    dqt = (U(i, j, k + 1) - U(i, j, k)) * dzc_inv[k];
    dqb = (U(i, j, k) - U(i, j, k - 1)) * dzc_inv[k - 1];
    conv = 0.25 * ((ue * ue - uw * uw) * dx_inv + (un * vn - us * vs) * dy_inv + (ut * wt - ub * wb) * dzf_inv[k]);
    conv = conv + 0.5 * u_crf * (ue - uw) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzf_inv[k]) * re_inv;
    *fu = visc - conv;

    ve = (V(i, j, k) + V(i + 1, j, k));
    vw = (V(i - 1, j, k) + V(i, j, k));
    vn = (V(i, j, k) + V(i, j + 1, k));
    vs = (V(i, j - 1, k) + V(i, j, k));
    vt = (V(i, j, k) * dzf[k + 1] + V(i, j, k + 1) * dzf[k]) * dzc_inv[k];
    vb = (V(i, j, k) * dzf[k - 1] + V(i, j, k - 1) * dzf[k]) * dzc_inv[k - 1];
    ue = (U(i, j, k) + U(i, j + 1, k));
    uw = (U(i - 1, j, k) + U(i - 1, j + 1, k));
    wt = (W(i, j, k) + W(i, j + 1, k));
    wb = (W(i, j, k - 1) + W(i, j + 1, k - 1));
    dqe = (V(i + 1, j, k) - V(i, j, k)) * dx_inv;
    dqw = (V(i, j, k) - V(i - 1, j, k)) * dx_inv;
    dqn = (V(i, j + 1, k) - V(i, j, k)) * dy_inv;
    dqs = (V(i, j, k) - V(i, j - 1, k)) * dy_inv;
    dqt = (V(i, j, k + 1) - V(i, j, k)) * dzc_inv[k];
    dqb = (V(i, j, k) - V(i, j, k - 1)) * dzc_inv[k - 1];
    conv = 0.25 * ((ue * ve - uw * vw) * dx_inv + (vn * vn - vs * vs) * dy_inv + (wt * vt - wb * vb) * dzf_inv[k]);
    conv = conv + 0.5 * u_crf * (ve - vw) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzf_inv[k]) * re_inv;
    *fv = visc - conv;

    we = (W(i, j, k) + W(i + 1, j, k));
    ww = (W(i - 1, j, k) + W(i, j, k));
    wn = (W(i, j, k) + W(i, j + 1, k));
    ws = (W(i, j - 1, k) + W(i, j, k));
    wt = (W(i, j, k) + W(i, j, k + 1));
    wb = (W(i, j, k) + W(i, j, k - 1));
    ue = (U(i, j, k) * dzf[k + 1] + U(i, j, k + 1) * dzf[k]) * dzc_inv[k];
    uw = (U(i - 1, j, k) * dzf[k + 1] + U(i - 1, j, k + 1) * dzf[k]) * dzc_inv[k];
    vn = (V(i, j, k) * dzf[k + 1] + V(i, j, k + 1) * dzf[k]) * dzc_inv[k];
    vs = (V(i, j - 1, k) * dzf[k + 1] + V(i, j - 1, k + 1) * dzf[k]) * dzc_inv[k];
    dqe = (W(i + 1, j, k) - W(i, j, k)) * dx_inv;
    dqw = (W(i, j, k) - W(i - 1, j, k)) * dx_inv;
    dqn = (W(i, j + 1, k) - W(i, j, k)) * dy_inv;
    dqs = (W(i, j, k) - W(i, j - 1, k)) * dy_inv;
    dqt = (W(i, j, k + 1) - W(i, j, k)) * dzf_inv[k + 1];
    dqb = (W(i, j, k) - W(i, j, k - 1)) * dzf_inv[k];
    conv = 0.25 * ((ue * we - uw * ww) * dx_inv + (vn * wn - vs * ws) * dy_inv + (wt * wt - wb * wb) * dzc_inv[k]);
    conv = conv + 0.5 * u_crf * (we - ww) * dx_inv;
    visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzc_inv[k]) * re_inv;
    *fw = visc - conv;
}
#undef U
#undef V
#undef W

/* b: NULL, or a body force (three haloed fields) added to the right-hand
 * side, f := (visc - conv) + b: an extension for the adjoint checks, not in
 * the reference. */
static void rk1_impl(const orc_grid_t *g, const double *u, const double *v, const double *w,
                     double *u1, double *v1, double *w1, double u_crf, const double *const *b)
{ /* mod_calcVel.f90:52-146: u1 = u + dt*(visc - conv) */
    const double dt = g->dt;
#pragma omp parallel for schedule(static) NT
    for (int k = 1; k <= g->nz; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                double fu, fv, fw;
                mom_rhs(g, u, v, w, i, j, k, u_crf, &fu, &fv, &fw);
                size_t e = IX(g, i, j, k);
                if (b) {
                    fu = fu + b[0][e];
                    fv = fv + b[1][e];
                    fw = fw + b[2][e];
                }
                u1[e] = u[e] + dt * fu;
                v1[e] = v[e] + dt * fv;
                w1[e] = w[e] + dt * fw;
            }
}

static void rk2_impl(const orc_grid_t *g, double *u, double *v, double *w, const double *u1,
                     const double *v1, const double *w1, double u_crf, const double *const *b)
{ /* mod_calcVel.f90:240-334: u = (u1 + dt*(visc - conv) + u)*0.5 */
    const double dt = g->dt;
#pragma omp parallel for schedule(static) NT
    for (int k = 1; k <= g->nz; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                double fu, fv, fw;
                mom_rhs(g, u1, v1, w1, i, j, k, u_crf, &fu, &fv, &fw);
                size_t e = IX(g, i, j, k);
                if (b) {
                    fu = fu + b[0][e];
                    fv = fv + b[1][e];
                    fw = fw + b[2][e];
                }
                u[e] = (u1[e] + dt * fu + u[e]) * 0.5;
                v[e] = (v1[e] + dt * fv + v[e]) * 0.5;
                w[e] = (w1[e] + dt * fw + w[e]) * 0.5;
            }
}

void orc_rk1(const orc_grid_t *g, const double *u, const double *v, const double *w,
             double *u1, double *v1, double *w1, double u_crf)
{
    rk1_impl(g, u, v, w, u1, v1, w1, u_crf, NULL);
}

void orc_rk2(const orc_grid_t *g, double *u, double *v, double *w,
             const double *u1, const double *v1, const double *w1, double u_crf)
{
    rk2_impl(g, u, v, w, u1, v1, w1, u_crf, NULL);
}

void orc_calc_rhs(const orc_grid_t *g, const double *u, const double *v,
                  const double *w, double *rhs)
{ /* mod_calcRHS.f90:11-37 */
    const double dti = 1.0 / g->dt;
    const double dtidxi = dti * g->dx_inv, dtidyi = dti * g->dy_inv;
#pragma omp parallel for schedule(static) NT
    for (int k = 1; k <= g->nz; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                rhs[IX(g, i, j, k)] = (u[IX(g, i, j, k)] - u[IX(g, i - 1, j, k)]) * dtidxi +
                                      (v[IX(g, i, j, k)] - v[IX(g, i, j - 1, k)]) * dtidyi +
                                      (w[IX(g, i, j, k)] - w[IX(g, i, j, k - 1)]) * dti * g->dzf_inv[k];
            }
}

void orc_correct_vel(const orc_grid_t *g, const double *p, double *u, double *v,
                     double *w)
{ /* mod_calcVel.f90:428-453 */
    const double dt = g->dt;
    const double dtdxi = dt * g->dx_inv, dtdyi = dt * g->dy_inv;
#pragma omp parallel for schedule(static) NT
    for (int k = 1; k <= g->nz; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                size_t e = IX(g, i, j, k);
                u[e] = u[e] - (p[IX(g, i + 1, j, k)] - p[e]) * dtdxi;
                v[e] = v[e] - (p[IX(g, i, j + 1, k)] - p[e]) * dtdyi;
                w[e] = w[e] - (p[IX(g, i, j, k + 1)] - p[e]) * dt * g->dzc_inv[k];
            }
}

double orc_mean_h(const orc_grid_t *g, const double *var)
{ /* mod_utils.f90:105-131; summed plane by plane, planes in k order */
    double *part = (double *)malloc(sizeof(double) * (size_t)g->nz);
#pragma omp parallel for schedule(static) NT
    for (int k = 1; k <= g->nz; k++) {
        double s = 0.0;
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) s = s + var[IX(g, i, j, k)] * g->dzflzi[k];
        part[k - 1] = s;
    }
    double m = 0.0;
    for (int k = 0; k < g->nz; k++) m += part[k];
    free(part);
    m = m / g->nx / g->ny;
    return m;
}

void orc_force_vel(const orc_grid_t *g, const int is_forced[3],
                   const double vel_force[3], double *u, double *v, double *w)
{ /* mod_calcVel.f90:455-509 */
    double *f3[3] = {u, v, w};
    for (int d = 0; d < 3; d++) {
        if (!is_forced[d]) continue;
        double mean = orc_mean_h(g, f3[d]);
        double force = vel_force[d] - mean;
        double *q = f3[d];
#pragma omp parallel for schedule(static) NT
        for (int k = 1; k <= g->nz; k++)
            for (int j = 1; j <= g->ny; j++)
                for (int i = 1; i <= g->nx; i++) q[IX(g, i, j, k)] = q[IX(g, i, j, k)] + force;
    }
}

/* updateHalo at one rank (mod_updateBound.f90:381-386): y periodic through
 * the self neighbour, full x and z extent; z exchange is a no-op at walls. */
static void halo_y(const orc_grid_t *g, double *q)
{
    for (int k = 0; k <= g->nz + 1; k++)
        for (int i = 0; i <= g->nx + 1; i++) q[IX(g, i, g->ny + 1, k)] = q[IX(g, i, 1, k)];
    for (int k = 0; k <= g->nz + 1; k++)
        for (int i = 0; i <= g->nx + 1; i++) q[IX(g, i, 0, k)] = q[IX(g, i, g->ny, k)];
}

/* imposePeriodicBC, idir = 1 (mod_updateBound.f90:241-272) */
static void periodic_x(const orc_grid_t *g, double *q)
{
    for (int k = 0; k <= g->nz + 1; k++)
        for (int j = 0; j <= g->ny + 1; j++) q[IX(g, 0, j, k)] = q[IX(g, g->nx, j, k)];
    for (int k = 0; k <= g->nz + 1; k++)
        for (int j = 0; j <= g->ny + 1; j++) q[IX(g, g->nx + 1, j, k)] = q[IX(g, 1, j, k)];
}

/* imposeNoSlipBC (mod_updateBound.f90:274-330) */
static void noslip(const orc_grid_t *g, int ibound, int centered, double *q, double bcvalue)
{
    const int n = g->nz;
    for (int j = 0; j <= g->ny + 1; j++)
        for (int i = 0; i <= g->nx + 1; i++) {
            if (ibound == 0) {
                if (centered) q[IX(g, i, j, 0)] = 2.0 * bcvalue - q[IX(g, i, j, 1)];
                else q[IX(g, i, j, 0)] = bcvalue;
            } else {
                if (centered) q[IX(g, i, j, n + 1)] = 2.0 * bcvalue - q[IX(g, i, j, n)];
                else { q[IX(g, i, j, n)] = bcvalue; q[IX(g, i, j, n + 1)] = bcvalue; }
            }
        }
}

void orc_bc_vel(const orc_grid_t *g, double *u, double *v, double *w, double u_crf)
{ /* updateBoundVel: mod_updateBound.f90:30-108 */
    halo_y(g, u); halo_y(g, v); halo_y(g, w);
    periodic_x(g, u); periodic_x(g, v); periodic_x(g, w);
    const double bcu = 0.0 - u_crf;
    noslip(g, 0, 1, u, bcu); noslip(g, 0, 1, v, 0.0); noslip(g, 0, 0, w, 0.0);
    noslip(g, 1, 1, u, bcu); noslip(g, 1, 1, v, 0.0); noslip(g, 1, 0, w, 0.0);
}

void orc_bc_p(const orc_grid_t *g, double *p)
{ /* updateBoundP: mod_updateBound.f90:186-239, imposeZeroGradBC :332-362 */
    halo_y(g, p);
    periodic_x(g, p);
    for (int j = 0; j <= g->ny + 1; j++)
        for (int i = 0; i <= g->nx + 1; i++) p[IX(g, i, j, 0)] = p[IX(g, i, j, 1)];
    for (int j = 0; j <= g->ny + 1; j++)
        for (int i = 0; i <= g->nx + 1; i++) p[IX(g, i, j, g->nz + 1)] = p[IX(g, i, j, g->nz)];
}

double orc_max_div(const orc_grid_t *g, const double *u, const double *v,
                   const double *w, int ijk[3])
{ /* mod_utils.f90:257-318: first location (k,j,i order) of the maximum */
    double best = 0.0;
    int bi = 1, bj = 1, bk = 1;
    for (int k = 1; k <= g->nz; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                double d = fabs((u[IX(g, i, j, k)] - u[IX(g, i - 1, j, k)]) * g->dx_inv +
                                (v[IX(g, i, j, k)] - v[IX(g, i, j - 1, k)]) * g->dy_inv +
                                (w[IX(g, i, j, k)] - w[IX(g, i, j, k - 1)]) * g->dzf_inv[k]);
                if (d > best) { best = d; bi = i; bj = j; bk = k; }
            }
    if (ijk) { ijk[0] = bi; ijk[1] = bj; ijk[2] = bk; }
    return best;
}

double orc_max_cfl(const orc_grid_t *g, const double *u, const double *v,
                   const double *w, int ijk[3])
{ /* mod_utils.f90:193-255: first location (k,j,i order) of the maximum, times dt */
    double best = 0.0;
    int bi = 1, bj = 1, bk = 1;
    for (int k = 1; k <= g->nz; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                double c = fabs((u[IX(g, i, j, k)] + u[IX(g, i - 1, j, k)]) * 0.5) * g->dx_inv +
                           fabs((v[IX(g, i, j, k)] + v[IX(g, i, j - 1, k)]) * 0.5) * g->dy_inv +
                           fabs((w[IX(g, i, j, k)] + w[IX(g, i, j, k - 1)]) * 0.5) * g->dzf_inv[k];
                if (c > best) { best = c; bi = i; bj = j; bk = k; }
            }
    if (ijk) { ijk[0] = bi; ijk[1] = bj; ijk[2] = bk; }
    return best * g->dt;
}

int orc_has_nan(long n, const double *q)
{ /* mod_utils.f90:352-394 over the whole array */
    for (long t = 0; t < n; t++)
        if (q[t] != q[t]) return 1;
    return 0;
}

void orc_wall_tau(const orc_grid_t *g, const double *u, double u_crf, double tau[2])
{ /* calcAndWriteSkf, mod_monitor.f90:225-258, summed in the reference's order
   * at the bottom (k = 1, dz = dzf(1)/2) and top (k = nz, dz = dzf(nz)/2)
   * wall (the calls at mod_monitor.f90:168-171) */
    const int kw[2] = {1, g->nz};
    for (int s = 0; s < 2; s++) {
        const double dz = 0.5 * g->dzf[kw[s]];
        double t = 0.0;
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) t = t + g->re_inv * (u[IX(g, i, j, kw[s])] + u_crf) / dz;
        tau[s] = t / g->nx / g->ny;
    }
}

double orc_wall_shear(const orc_grid_t *g, const double *u, double u_crf)
{ /* the two walls' tau_w averaged */
    double tau[2];
    orc_wall_tau(g, u, u_crf, tau);
    return 0.5 * (tau[0] + tau[1]);
}

void orc_step_body(const orc_grid_t *g, orc_ppe_t *ppe, double *u, double *v, double *w,
                   double *u1, double *v1, double *w1, double *p, double u_crf,
                   const int is_forced[3], const double vel_force[3], const double *bu,
                   const double *bv, const double *bw)
{ /* orc_step with a body force in both RK stages (not in the reference) */
    const double *b[3] = {bu, bv, bw};
    rk1_impl(g, u, v, w, u1, v1, w1, u_crf, b);
    orc_bc_vel(g, u1, v1, w1, u_crf);
    rk2_impl(g, u, v, w, u1, v1, w1, u_crf, b);
    orc_bc_vel(g, u, v, w, u_crf);
    orc_calc_rhs(g, u, v, w, p);
    orc_ppe_solve(ppe, p);
    orc_bc_p(g, p);
    orc_correct_vel(g, p, u, v, w);
    orc_force_vel(g, is_forced, vel_force, u, v, w);
    orc_bc_vel(g, u, v, w, u_crf);
}

/* orc_step with the wall time of each of the reference's GPTL buckets of the
 * main loop (src/main.f90:131-212): t = {uvw1, Update boundary vel (three
 * calls summed), uvw2, Calculate RHS, Poisson solver, Update boundary pres,
 * Correct vel, Force vel}, in seconds. */
void orc_step_timed(const orc_grid_t *g, orc_ppe_t *ppe, double *u, double *v, double *w,
                    double *u1, double *v1, double *w1, double *p, double u_crf,
                    const int is_forced[3], const double vel_force[3], double t[8])
{
    double c0 = now_s(), c1;
    for (int i = 0; i < 8; i++) t[i] = 0.0;
#define LAP(slot) (c1 = now_s(), t[slot] += c1 - c0, c0 = c1)
    orc_rk1(g, u, v, w, u1, v1, w1, u_crf);
    LAP(0);
    orc_bc_vel(g, u1, v1, w1, u_crf);
    LAP(1);
    orc_rk2(g, u, v, w, u1, v1, w1, u_crf);
    LAP(2);
    orc_bc_vel(g, u, v, w, u_crf);
    LAP(1);
    orc_calc_rhs(g, u, v, w, p);
    LAP(3);
    orc_ppe_solve(ppe, p);
    LAP(4);
    orc_bc_p(g, p);
    LAP(5);
    orc_correct_vel(g, p, u, v, w);
    LAP(6);
    orc_force_vel(g, is_forced, vel_force, u, v, w);
    LAP(7);
    orc_bc_vel(g, u, v, w, u_crf);
    LAP(1);
#undef LAP
}

void orc_step(const orc_grid_t *g, orc_ppe_t *ppe, double *u, double *v, double *w,
              double *u1, double *v1, double *w1, double *p, double u_crf,
              const int is_forced[3], const double vel_force[3])
{ /* src/main.f90:137-208 without NB_HALO */
    orc_rk1(g, u, v, w, u1, v1, w1, u_crf);
    orc_bc_vel(g, u1, v1, w1, u_crf);
    orc_rk2(g, u, v, w, u1, v1, w1, u_crf);
    orc_bc_vel(g, u, v, w, u_crf);
    orc_calc_rhs(g, u, v, w, p);
    orc_ppe_solve(ppe, p);
    orc_bc_p(g, p);
    orc_correct_vel(g, p, u, v, w);
    orc_force_vel(g, is_forced, vel_force, u, v, w);
    orc_bc_vel(g, u, v, w, u_crf);
}
