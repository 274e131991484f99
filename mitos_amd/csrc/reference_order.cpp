// The reference's own call order through the drop-in functions, in FP64:
// src/mitos.c:148,216-227,300,309-317,367-416, src/generate_grids.c:52-60,
// src/mitos.c:794-863,1072-1077 and analyse/src/mitos_read.c:278-419.
// Every transform and kernel is a separate host-synchronous call, as in the
// reference, so this is the "plain FP64" GPU path the fused pipeline is
// measured against.
#include <hip/hip_runtime.h>
#include <string.h>
#include <time.h>

#include <vector>

#include "../../include/mitos_hip.h"
#include "runtime.h"

namespace {

double now() {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

struct Grids {
    std::vector<struct distributed_grid*> all;
    ~Grids() {
        for (auto* g : all) {
            free_local_grid(g);
            delete g;
        }
    }
    struct distributed_grid* make(int N, double L, int* rc) {
        auto* g = new struct distributed_grid;
        memset(g, 0, sizeof(*g));
        all.push_back(g);
        *rc = alloc_local_grid(g, N, L, MPI_COMM_WORLD);
        return g;
    }
    void drop(struct distributed_grid* g) {
        free_local_grid(g);
    }
};

}  // namespace

#define RTRY(expr)                       \
    do {                                 \
        int _rc = (expr);                \
        if (_rc != MITOS_OK) return _rc; \
    } while (0)

extern "C" int mitos_ic_run_reference_order(const struct mitos_ic_config* cfg,
                                            struct mitos_ic_result* res, float* X, float* Y,
                                            float* Z) {
    if (!cfg || cfg->precision != MITOS_FP64) return MITOS_ERR_ARG;
    const int N = cfg->N;
    const double L = cfg->boxlen;
    const long long np = (long long)N * N * N;
    // particle lattice M^3 (src/particle.c:87-129; M = N unless cfg->M is set)
    const int Ml = cfg->M > 0 ? cfg->M : N;
    const long long npart = (long long)Ml * Ml * Ml;
    double t0 = now(), t1;
    float stage[8] = {0};
    int rc = MITOS_OK;
    Grids G;

    const double A = mitos_sigma8_amplitude(cfg->n_s, cfg->gamma, cfg->sigma8, 8.0);
    struct mitos_spectrum_params sp;
    sp.A = A;
    sp.n_s = cfg->n_s;
    sp.Gamma = cfg->gamma;

    // src/mitos.c:148, 216-227, 300
    rng_state seed = rand_uint64_init(cfg->seed);
    struct distributed_grid* grf = G.make(N, L, &rc);
    RTRY(rc);
    RTRY(generate_complex_grf(grf, &seed));
    RTRY(enforce_hermiticity(grf));
    RTRY(fft_apply_kernel_dg(grf, grf, kernel_sqrt_power_law, &sp));
    RTRY(fft_c2r_dg(grf));
    RTRY(fft_r2c_dg(grf));
    t1 = now();
    stage[0] = (float)(1e3 * (t1 - t0));
    t0 = t1;

    // src/mitos.c:309-317, src/generate_grids.c:52-60, src/mitos.c:377
    struct distributed_grid* grid = G.make(N, L, &rc);
    RTRY(rc);
    struct distributed_grid* potential = G.make(N, L, &rc);
    RTRY(rc);
    struct distributed_grid* derivative = G.make(N, L, &rc);
    RTRY(rc);
    RTRY(fft_apply_kernel_dg(grid, grf, kernel_bbks_transfer, &sp));
    RTRY(fft_c2r_dg(grid));
    G.drop(grf);
    RTRY(fft_r2c_dg(grid));

    // src/mitos.c:380-391
    if (cfg->run2lpt) {
        RTRY(solve2LPT(potential, grid, derivative, cfg->f1, cfg->f2));
    } else {
        RTRY(fft_apply_kernel_dg(potential, grid, kernel_inv_poisson, NULL));
    }
    G.drop(grid);
    t1 = now();
    stage[1] = (float)(1e3 * (t1 - t0));
    t0 = t1;

    // src/mitos.c:394-416: undo TSC window, three derivative grids
    struct Hermite_kern_params hk;
    hk.order = 3;
    hk.N = N;
    hk.boxlen = L;
    RTRY(fft_apply_kernel_dg(potential, potential, kernel_undo_Hermite_window, &hk));
    void (*const dker[3])(struct kernel*) = {kernel_dx, kernel_dy, kernel_dz};
    struct distributed_grid* psi[3];
    for (int i = 0; i < 3; i++) {
        RTRY(fft_apply_kernel_dg(derivative, potential, dker[i], NULL));
        RTRY(fft_c2r_dg(derivative));
        psi[i] = G.make(N, L, &rc);
        RTRY(rc);
        free_local_complex_grid(psi[i]);
        // on the library stream, ahead of the next derivative into the same grid
        if (hipMemcpyAsync(psi[i]->box, derivative->box, sizeof(double) * N * (long)N * (N + 2),
                           hipMemcpyDeviceToDevice, mitos::api_stream()) != hipSuccess ||
            mitos::api_sync() != MITOS_OK)
            return MITOS_ERR_DEVICE;
        psi[i]->momentum_space = 0;
    }
    G.drop(potential);
    G.drop(derivative);
    t1 = now();
    stage[3] = (float)(1e3 * (t1 - t0));
    t0 = t1;

    // src/mitos.c:794-863, 1072-1077
    float *dX = X, *dY = Y, *dZ = Z;
    std::vector<void*> tmp;
    if (!dX || !dY || !dZ) {
        dX = (float*)mitos::pool_alloc(sizeof(float) * npart);
        dY = (float*)mitos::pool_alloc(sizeof(float) * npart);
        dZ = (float*)mitos::pool_alloc(sizeof(float) * npart);
        tmp = {dX, dY, dZ};
        if (!dX || !dY || !dZ) {
            for (void* p : tmp) mitos::pool_free(p);
            return MITOS_ERR_ALLOC;
        }
    }
    rc = mitos_lattice(Ml, L, 1, dX, dY, dZ, NULL);
    if (rc == MITOS_OK) rc = mitos_displace_tsc_dg(psi[0], psi[1], psi[2], npart, dX, dY, dZ);
    for (int i = 0; i < 3; i++) G.drop(psi[i]);
    t1 = now();
    stage[4] = (float)(1e3 * (t1 - t0));
    t0 = t1;

    // analyse/src/mitos_read.c:278-419
    double* box = nullptr;
    fftw_complex* fbox = nullptr;
    if (rc == MITOS_OK) {
        box = (double*)mitos::pool_alloc(sizeof(double) * np);
        fbox = (fftw_complex*)mitos::pool_alloc(sizeof(double) * 2 * N * (long)N * (N / 2 + 1));
        if (!box || !fbox) rc = MITOS_ERR_ALLOC;
    }
    if (rc == MITOS_OK) rc = mitos_cic_overdensity(N, L, npart, dX, dY, dZ, box);
    t1 = now();
    stage[5] = (float)(1e3 * (t1 - t0));
    t0 = t1;
    std::vector<double> k(cfg->bins), P(cfg->bins);
    std::vector<int> obs(cfg->bins);
    if (rc == MITOS_OK) rc = mitos_fft_r2c(N, box, fbox);
    if (rc == MITOS_OK) rc = fft_normalize_r2c(fbox, N, L);
    if (rc == MITOS_OK) rc = undoCICWindow(fbox, N, L);
    if (rc == MITOS_OK) {
        calc_cross_powerspec(N, L, fbox, fbox, cfg->bins, k.data(), P.data(), obs.data());
        rc = mitos_last_error();
    }
    t1 = now();
    stage[6] = (float)(1e3 * (t1 - t0));
    for (void* p : tmp) mitos::pool_free(p);
    mitos::pool_free(box);
    mitos::pool_free(fbox);
    if (rc != MITOS_OK) return rc;
    if (res) {
        for (int i = 0; i < cfg->bins; i++) {
            if (res->k_in_bins) res->k_in_bins[i] = k[i];
            if (res->power_in_bins) res->power_in_bins[i] = P[i];
            if (res->obs_in_bins) res->obs_in_bins[i] = obs[i];
        }
        res->amplitude = A;
        float tot = 0;
        for (int i = 0; i < 7; i++) {
            res->stage_ms[i] = stage[i];
            tot += stage[i];
        }
        res->stage_ms[7] = tot;
    }
    return MITOS_OK;
}
