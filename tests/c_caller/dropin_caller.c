/*
 * A compiled caller of the drop-in, in the order the reference's Fortran main
 * calls the C solver through ISO_C_BINDING (src/mod_poissonSolver.f90:59-85,
 * src/main.f90:177): init_poisson_solver once, execute_poisson_solver on the
 * haloed host array p(0:nx+1,0:ny+1,0:nz+1) each step, free_poisson_solver at
 * the end.  Nothing here knows about the GPU.
 *
 *   dropin_caller IN OUT
 * IN:  int32 nx, ny, nz; float64 dx, dy; float64 dzf[nz+2];
 *      float64 var[(nz+2)(ny+2)(nx+2)] (x fastest)
 * OUT: float64 var after the first call, then after a second call made on a
 *      fresh copy of the input (the same answer is expected both times).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "powerllel.h"

static int rd(void *p, size_t sz, size_t n, FILE *f) { return fread(p, sz, n, f) == n ? 0 : -1; }

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 3;
    int dims[3];
    double dxy[2];
    if (rd(dims, sizeof(int), 3, f) || rd(dxy, sizeof(double), 2, f)) return 4;
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    const size_t n = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    double *dzf = malloc(sizeof(double) * (size_t)(nz + 2));
    double *rhs = malloc(sizeof(double) * n);
    double *var = malloc(sizeof(double) * n);
    if (!dzf || !rhs || !var) return 5;
    if (rd(dzf, sizeof(double), (size_t)nz + 2, f) || rd(rhs, sizeof(double), n, f)) return 6;
    fclose(f);

    /* one rank: x and y periodic on the rank itself, walls in z */
    int neighbor_xyz[3][6] = {{0}};
    neighbor_xyz[2][4] = -1;
    neighbor_xyz[2][5] = -1;
    init_poisson_solver(nx, ny, nz, dxy[0], dxy[1], dzf, "PP", "PP", "NN", neighbor_xyz);

    FILE *o = fopen(argv[2], "wb");
    if (!o) return 7;
    for (int call = 0; call < 2; call++) {
        memcpy(var, rhs, sizeof(double) * n);
        execute_poisson_solver(var);
        if (fwrite(var, sizeof(double), n, o) != n) return 8;
    }
    fclose(o);
    free_poisson_solver();
    free(dzf);
    free(rhs);
    free(var);
    return 0;
}
