/*
 * A C caller of libmitos_hip.so written the way a reference translation unit
 * calls it (src/mitos.c:216-300, src/2lpt.c:75-136): struct distributed_grid
 * grids from alloc_local_grid, the seed stream, generate_complex_grf, and
 * k-space kernels passed to fft_apply_kernel_dg both as the library's
 * exported callbacks (include/fft_kernels.h of this library: device kernels)
 * and as static inline copies local to this unit, the form the reference's
 * own include/fft_kernels.h gives every unit (evaluated on the host by the
 * library, mode by mode).  Both must give the same grid.
 *
 * Prints one line per check ("name max_rel_diff") and exits 0 when every
 * check is within its bound, 1 otherwise, 2 on a library error.
 *
 * Build: gcc -std=gnu11 -O2 -I include tests/c_caller/dropin_caller.c \
 *            -L mitos_amd -lmitos_hip -Wl,-rpath,<abs path of mitos_amd> -lm
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mitos_hip.h"
#include "fft_kernels.h"

/* unit-local callbacks, as static inline definitions in a header give them */
static inline void local_kernel_dx(struct kernel *the_kernel) {
    the_kernel->kern = I * the_kernel->kx;
}
static inline void local_kernel_inv_poisson(struct kernel *the_kernel) {
    const double k = the_kernel->k;
    the_kernel->kern = (k > 0) ? -1.0 / k / k : 1.0;
}
static inline void local_kernel_gaussian(struct kernel *the_kernel) {
    const double kR = the_kernel->k * *((const double *)the_kernel->params);
    the_kernel->kern = exp(-kR * kR);
}
/* a kernel of the caller's own, with no device counterpart */
static inline void local_kernel_shell(struct kernel *the_kernel) {
    const double *p = (const double *)the_kernel->params;
    the_kernel->kern = (the_kernel->k > p[0] && the_kernel->k < p[1]) ? 1.0 + I * 0.5 : 0.0;
}

static double max_rel(const fftw_complex *a, const fftw_complex *b, long n) {
    double num = 0, den = 0;
    for (long i = 0; i < n; i++) {
        const double d = cabs(a[i] - b[i]);
        if (d > num) num = d;
        if (cabs(b[i]) > den) den = cabs(b[i]);
    }
    return den > 0 ? num / den : num;
}

#define CHECK(call)                                                                    \
    do {                                                                               \
        int _e = (call);                                                               \
        if (_e) {                                                                      \
            fprintf(stderr, "%s failed: %d (%s)\n", #call, _e, mitos_strerror(_e));    \
            return 2;                                                                  \
        }                                                                              \
    } while (0)

int main(void) {
    const int N = 32;
    const double boxlen = 150.0;
    const long n = (long)N * N * (N / 2 + 1);
    struct distributed_grid grf, a, b;
    CHECK(alloc_local_grid(&grf, N, boxlen, MPI_COMM_WORLD));
    CHECK(alloc_local_grid(&a, N, boxlen, MPI_COMM_WORLD));
    CHECK(alloc_local_grid(&b, N, boxlen, MPI_COMM_WORLD));
    rng_state seed = rand_uint64_init(42);
    CHECK(generate_complex_grf(&grf, &seed));
    CHECK(enforce_hermiticity(&grf));

    fftw_complex *ha = malloc(sizeof(fftw_complex) * n);
    fftw_complex *hb = malloc(sizeof(fftw_complex) * n);
    int bad = 0;

    /* exported (device) against unit-local (host-evaluated) callbacks */
    const double R = 3.5;
    struct {
        const char *name;
        kernel_func lib, local;
        const void *params;
        double bound;
    } cases[] = {
        {"kernel_dx", kernel_dx, local_kernel_dx, NULL, 0.0},
        {"kernel_inv_poisson", kernel_inv_poisson, local_kernel_inv_poisson, NULL, 0.0},
        {"kernel_gaussian", kernel_gaussian, local_kernel_gaussian, &R, 1e-15},
    };
    for (int c = 0; c < 3; c++) {
        CHECK(fft_apply_kernel_dg(&a, &grf, cases[c].lib, cases[c].params));
        CHECK(fft_apply_kernel_dg(&b, &grf, cases[c].local, cases[c].params));
        CHECK(mitos_grid_to_host(&a, ha, NULL));
        CHECK(mitos_grid_to_host(&b, hb, NULL));
        const double d = max_rel(hb, ha, n);
        printf("%s %.3e\n", cases[c].name, d);
        if (!(d <= cases[c].bound)) bad = 1;
    }

    /* a caller-defined kernel: the host-evaluated multiplier against the
     * same product formed here on the host */
    const double shell[2] = {0.2, 0.6};
    CHECK(fft_apply_kernel_dg(&a, &grf, local_kernel_shell, shell));
    CHECK(mitos_grid_to_host(&a, ha, NULL));
    CHECK(mitos_grid_to_host(&grf, hb, NULL));
    const double dk = 2 * M_PI / boxlen;
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++)
            for (int z = 0; z <= N / 2; z++) {
                double kx, ky, kz, k;
                fft_wavevector(x, y, z, N, dk, &kx, &ky, &kz, &k);
                struct kernel kk = {kx, ky, kz, k, 0, shell};
                local_kernel_shell(&kk);
                const long id = ((long)x * N + y) * (N / 2 + 1) + z;
                hb[id] = hb[id] * kk.kern;
            }
    const double ds = max_rel(ha, hb, n);
    printf("caller_kernel %.3e\n", ds);
    if (!(ds == 0.0)) bad = 1;

    /* 2LPT through the reference's call (src/mitos.c:385): the momentum-space
     * check of src/2lpt.c:69-72 returns 1 */
    struct distributed_grid pot;
    CHECK(alloc_local_grid(&pot, N, boxlen, MPI_COMM_WORLD));
    a.momentum_space = 0;
    const int e = solve2LPT(&pot, &a, &b, 1.0, -3. / 7.);
    printf("solve2LPT_not_momentum %d\n", e);
    if (e != 1) bad = 1;

    free(ha);
    free(hb);
    free_local_grid(&pot);
    free_local_grid(&grf);
    free_local_grid(&a);
    free_local_grid(&b);
    printf(bad ? "FAIL\n" : "OK\n");
    return bad;
}
