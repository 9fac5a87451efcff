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

#include "pll_internal.h"

namespace {

typedef double2 cd;

struct TArgs {
    const double *a, *c, *piv;
    int L, ny, nz, pin_gauge;
};

__global__ void __launch_bounds__(256) k_trid(cd *__restrict__ W, TArgs g)
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
    cd *__restrict__ col = W + s;

    cd prev = col[0];
    for (int k = 1; k < nz; ++k) {
        cd cur = col[(long long)k * nslot];
        const double mr = a[k] / pr[(long long)(k - 1) * pst];
        const double mi = same ? mr : a[k] / pi[(long long)(k - 1) * pst];
        cur.x -= mr * prev.x;
        cur.y -= mi * prev.y;
        col[(long long)k * nslot] = cur;
        prev = cur;
    }
    {
        const double br = pr[(long long)(nz - 1) * pst];
        const double bi = pi[(long long)(nz - 1) * pst];
        const bool pin = g.pin_gauge && s == 0;
        prev.x = pin ? 0.0 : (br != 0.0 ? prev.x / br : 0.0);
        prev.y = bi != 0.0 ? prev.y / bi : 0.0;
        col[(long long)(nz - 1) * nslot] = prev;
    }
    for (int k = nz - 2; k >= 0; --k) {
        cd cur = col[(long long)k * nslot];
        const double br = pr[(long long)k * pst];
        const double bi = same ? br : pi[(long long)k * pst];
        cur.x = (cur.x - c[k] * prev.x) / br;
        cur.y = (cur.y - c[k] * prev.y) / bi;
        col[(long long)k * nslot] = cur;
        prev = cur;
    }
}

}  // namespace

extern "C" int pll_launch_trid(const pll_ppe_dev *d, void *W, hipStream_t s)
{
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
    hipLaunchKernelGGL(k_trid, dim3((unsigned)blocks), dim3(256), 0, s, (cd *)W, g);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
