// calcMaxDiv without any device allocation (pll_max_div_noalloc).
//
// For callers that cannot call hipMalloc (code under stream capture, tight
// loops, hosts that forbid device allocation): the per-row partial maxima
// that pll_channel_max_div keeps in its channel handle live here in one
// file-scope device array per GPU, `scratch`.  Every call on a GPU shares
// that GPU's array; calls that overlap on one GPU race.  Each call
// overwrites every part of the array it reads (the 1/dzf table, one partial
// per (j,k) row, the final pair), so nothing needs clearing between calls.
//
// The reduction kernels are those of pll_channel_max_div (step_kernels.hip,
// src/mod_utils.f90:257-318), so the two entry points give identical results.
#include <hip/hip_runtime.h>

#include <pthread.h>

#include "pll_internal.h"
#include "powerllel.h"

extern "C" int pll_set_error(int code, const char *fmt, ...);

namespace {
constexpr size_t kZB = sizeof(double) * PLL_NOALLOC_MAX_NZ2;
constexpr size_t kRB = sizeof(double) * PLL_NOALLOC_MAX_ROWS;
constexpr size_t kBytes = kZB + 2 * kRB + 256;
}  // namespace

// The work space of every call on this GPU (layout below).
static __device__ __align__(256) unsigned char scratch[kBytes];

namespace {

// Layout of the scratch array (each region 256-byte aligned):
//   [0, ZB)              1/dzf_global[k], k = 0..nz+1
//   [ZB, ZB + RB)        partial maxima, one double per (j,k) row
//   [ZB + RB, ZB + 2RB)  their linear indices, one int64 per row
//   [ZB + 2RB, +256)     final maximum (double), then at +128 its index (int64)
static_assert(kZB % 256 == 0 && kRB % 256 == 0, "regions stay 256-byte aligned");
static_assert(kBytes == PLL_NOALLOC_SCRATCH_BYTES, "header and layout agree");

constexpr int kMaxDevices = 64;
pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
unsigned char *g_addr[kMaxDevices];  // device address of `scratch`, per device

// The device address of this GPU's scratch array, looked up once per device
// with that device current.  The host symbol `scratch` is a stand-in, never a
// device address: kernels and copies take the looked-up pointer.
int scratch_addr(unsigned char **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return pll_set_error(PLL_EHIP, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return pll_set_error(PLL_EUNSUP, "device %d beyond the %d the scratch table holds", dev,
                             kMaxDevices);
    pthread_mutex_lock(&g_mu);
    if (!g_addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            pthread_mutex_unlock(&g_mu);
            return pll_set_error(PLL_EHIP, "hipGetSymbolAddress(scratch) failed: %s",
                                 hipGetErrorString(e));
        }
        g_addr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_addr[dev];
    pthread_mutex_unlock(&g_mu);
    return PLL_OK;
}

// inv[k] = 1 / dzf[k] (correctly rounded, as the channel's host table)
__global__ void __launch_bounds__(256) k_inv(const double *__restrict__ dzf, double *__restrict__ inv,
                                             int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) inv[k] = 1.0 / dzf[k];
}

}  // namespace

extern "C" int pll_max_div_noalloc(const double *u, const double *v, const double *w, int nx,
                                   int ny, int nz, double dx, double dy, const double *dzf_global,
                                   double *value, int ijk[3], void *stream)
{
    if (!u || !v || !w || !dzf_global) return pll_set_error(PLL_EINVAL, "NULL array");
    if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0.0) || !(dy > 0.0))
        return pll_set_error(PLL_EINVAL, "bad grid (%d %d %d, dx %g, dy %g)", nx, ny, nz, dx, dy);
    const long long rows = (long long)ny * nz;
    if (rows > PLL_NOALLOC_MAX_ROWS || nz + 2 > PLL_NOALLOC_MAX_NZ2)
        return pll_set_error(PLL_EINVAL,
                             "grid %dx%dx%d beyond the scratch array: ny*nz <= %d and nz <= %d",
                             nx, ny, nz, PLL_NOALLOC_MAX_ROWS, PLL_NOALLOC_MAX_NZ2 - 2);
    unsigned char *base = nullptr;
    int rc = scratch_addr(&base);
    if (rc) return rc;
    double *inv = reinterpret_cast<double *>(base);
    double *pval = reinterpret_cast<double *>(base + kZB);
    long long *pidx = reinterpret_cast<long long *>(base + kZB + kRB);
    double *oval = reinterpret_cast<double *>(base + kZB + 2 * kRB);
    long long *oidx = reinterpret_cast<long long *>(base + kZB + 2 * kRB + 128);
    hipStream_t s = static_cast<hipStream_t>(stream);

    const int n2 = nz + 2;
    hipLaunchKernelGGL(k_inv, dim3((n2 + 255) / 256), dim3(256), 0, s, dzf_global, inv, n2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pll_set_error(PLL_EHIP, "1/dzf launch failed: %s", hipGetErrorString(e));
    pll_grid_dev g = {};
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.dx_inv = 1.0 / dx;
    g.dy_inv = 1.0 / dy;
    g.dzf = dzf_global;
    g.dzf_inv = inv;
    if (pll_launch_maxdiv_partials(&g, u, v, w, pval, pidx, s) ||
        pll_launch_maxdiv_final((int)rows, pval, pidx, oval, oidx, s))
        return pll_set_error(PLL_EHIP, "max-div launch failed: %s",
                             hipGetErrorString(hipGetLastError()));
    double val;
    long long idx;
    if ((e = hipMemcpyAsync(&val, oval, sizeof val, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(&idx, oidx, sizeof idx, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return pll_set_error(PLL_EHIP, "max-div read-back failed: %s", hipGetErrorString(e));
    if (value) *value = val < 0.0 ? 0.0 : val;
    if (ijk) {
        ijk[0] = (int)(idx % nx) + 1;
        ijk[1] = (int)((idx / nx) % ny) + 1;
        ijk[2] = (int)(idx / ((long long)nx * ny)) + 1;
    }
    return PLL_OK;
}
