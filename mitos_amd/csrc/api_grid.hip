// Drop-in grid allocation, RNG and GRF entry points
// (include/distributed_grid.h:74-77, include/random.h:27-33, include/grf.h:30-31).
#include <map>
#include <mutex>

#include "../../include/mitos_hip.h"
#include "api_util.h"
#include "grf.h"
#include "runtime.h"

using namespace mitos;

namespace mitos {

thread_local int g_last_error = MITOS_OK;
int set_last_error(int e) {
    g_last_error = e;
    return e;
}

// ---- noise jump tables -------------------------------------------------------
struct JumpEntry {
    uint64_t* dev;
    int nbits;
};

int noise_jumps(int N, const Xo& s0, NoiseJumps* out) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, JumpEntry> cache;
    int dev = 0;
    MITOS_HIP_CHECK(hipGetDevice(&dev));
    const long H = N / 2 + 1;
    const long rows = (long)N * N;
    JumpEntry ent{nullptr, 0};
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(std::make_pair(dev, N));
        if (it != cache.end()) ent = it->second;
    }
    if (!ent.dev) {
        int nbits = 0;
        while (nbits < 62 && (1L << nbits) < rows) nbits++;
        if (nbits == 0) nbits = 1;
        std::vector<uint64_t> tab((size_t)nbits * 1024);
        BitMat J;
        bm_pow_step((uint64_t)(4 * H), J);
        for (int b = 0; b < nbits; b++) {
            memcpy(&tab[(size_t)b * 1024], J.rows, sizeof(J.rows));
            if (b + 1 < nbits) bm_mul(J, J, J);
        }
        uint64_t* d = nullptr;
        MITOS_HIP_CHECK(hipMalloc(&d, tab.size() * sizeof(uint64_t)));
        MITOS_HIP_CHECK(copy_sync(d, tab.data(), tab.size() * sizeof(uint64_t),
                                  hipMemcpyHostToDevice));
        ent = JumpEntry{d, nbits};
        std::lock_guard<std::mutex> lock(mu);
        cache[std::make_pair(dev, N)] = ent;
    }
    out->tab = ent.dev;
    out->nbits = ent.nbits;
    out->s0 = s0;
    Xo b = s0;
    for (long i = 0; i < 4 * H - 4; i++) xo_next(b);
    out->base = b;
    return MITOS_OK;
}

int noise_advance(int N, Xo* st) {
    static std::mutex mu;
    static std::map<int, BitMat*> cache;
    const uint64_t H = N / 2 + 1;
    const uint64_t draws = 4 * ((uint64_t)N * N * H - 1);
    BitMat* m = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(N);
        if (it != cache.end()) m = it->second;
    }
    if (!m) {
        m = new BitMat;
        bm_pow_step(draws, *m);
        std::lock_guard<std::mutex> lock(mu);
        cache[N] = m;
    }
    *st = bm_apply_rows(&m->rows[0][0], *st);
    return MITOS_OK;
}

}  // namespace mitos

// ---- Part 1: grids -------------------------------------------------------------
extern "C" int alloc_local_grid(struct distributed_grid* dg, int N, double boxlen,
                                MPI_Comm comm) {
    if (!dg || N < 2 || (N & 1)) return set_last_error(MITOS_ERR_ARG);
    const long local = (long)N * N * (N / 2 + 1);
    dg->N = N;
    dg->boxlen = boxlen;
    dg->comm = comm;
    dg->momentum_space = 0;
    dg->NX = N;
    dg->X0 = 0;
    dg->local_size = local;
    dg->fbox = nullptr;
    dg->box = nullptr;
    void* f = pool_alloc(sizeof(cx<double>) * local);
    if (!f) return set_last_error(MITOS_ERR_ALLOC);
    void* b = pool_alloc(sizeof(double) * 2 * local);
    if (!b) {
        pool_free(f);
        return set_last_error(MITOS_ERR_ALLOC);
    }
    // zero-filled, so padding entries are defined; on the library stream and
    // finished before return (the kernels of later calls run on that
    // non-blocking stream, which a null-stream memset would not order)
    if (hipMemsetAsync(f, 0, sizeof(cx<double>) * local, api_stream()) != hipSuccess ||
        hipMemsetAsync(b, 0, sizeof(double) * 2 * local, api_stream()) != hipSuccess ||
        api_sync() != MITOS_OK) {
        pool_free(f);
        pool_free(b);
        return set_last_error(MITOS_ERR_DEVICE);
    }
    dg->fbox = reinterpret_cast<fftw_complex*>(f);
    dg->box = reinterpret_cast<double*>(b);
    return MITOS_OK;
}

extern "C" int free_local_real_grid(struct distributed_grid* dg) {
    if (!dg) return MITOS_ERR_ARG;
    if (dg->box) pool_free(dg->box);
    dg->box = nullptr;
    return MITOS_OK;
}

extern "C" int free_local_complex_grid(struct distributed_grid* dg) {
    if (!dg) return MITOS_ERR_ARG;
    if (dg->fbox) pool_free(dg->fbox);
    dg->fbox = nullptr;
    return MITOS_OK;
}

extern "C" int free_local_grid(struct distributed_grid* dg) {
    free_local_real_grid(dg);
    return free_local_complex_grid(dg);
}

extern "C" int mitos_grid_to_host(const struct distributed_grid* dg, fftw_complex* fbox_host,
                                  double* box_host) {
    if (!dg) return set_last_error(MITOS_ERR_ARG);
    if (fbox_host && dg->fbox)
        MITOS_HIP_CHECK(copy_sync(fbox_host, dg->fbox, sizeof(cx<double>) * dg->local_size,
                                  hipMemcpyDeviceToHost));
    if (box_host && dg->box)
        MITOS_HIP_CHECK(copy_sync(box_host, dg->box, sizeof(double) * 2 * dg->local_size,
                                  hipMemcpyDeviceToHost));
    return MITOS_OK;
}

extern "C" int mitos_grid_from_host(struct distributed_grid* dg, const fftw_complex* fbox_host,
                                    const double* box_host) {
    if (!dg) return set_last_error(MITOS_ERR_ARG);
    if (fbox_host && dg->fbox)
        MITOS_HIP_CHECK(copy_sync(dg->fbox, fbox_host, sizeof(cx<double>) * dg->local_size,
                                  hipMemcpyHostToDevice));
    if (box_host && dg->box)
        MITOS_HIP_CHECK(copy_sync(dg->box, box_host, sizeof(double) * 2 * dg->local_size,
                                  hipMemcpyHostToDevice));
    return MITOS_OK;
}

// ---- Part 1: RNG ------------------------------------------------------------------
extern "C" uint64_t rand_uint64(rng_state* state) {
    Xo s;
    memcpy(s.s, state->s, sizeof(s.s));
    const uint64_t r = xo_next(s);
    memcpy(state->s, s.s, sizeof(s.s));
    return r;
}

extern "C" rng_state rand_uint64_init(uint64_t seed) {
    Xo s = xo_seed(seed);
    rng_state r;
    memcpy(r.s, s.s, sizeof(r.s));
    return r;
}

// ---- Part 1: GRF ------------------------------------------------------------------
extern "C" int generate_complex_grf(struct distributed_grid* dg, rng_state* state) {
    if (!dg || !state || !dg->fbox) return set_last_error(MITOS_ERR_ARG);
    const int N = dg->N;
    Xo s0;
    memcpy(s0.s, state->s, sizeof(s0.s));
    NoiseJumps J;
    MITOS_TRY(set_last_error(noise_jumps(N, s0, &J)));
    hipStream_t s = api_stream();
    MITOS_TRY(set_last_error(launch_grf<double, AmpOne>(reinterpret_cast<cx<double>*>(dg->fbox),
                                                        N / 2 + 1, N, dg->boxlen, 1, 0, J,
                                                        AmpOne(), s)));
    Xo s1 = s0;
    noise_advance(N, &s1);
    memcpy(state->s, s1.s, sizeof(s1.s));
    dg->momentum_space = 1;
    return set_last_error(api_sync());
}

extern "C" int enforce_hermiticity(struct distributed_grid* dg) {
    if (!dg || !dg->fbox) return set_last_error(MITOS_ERR_ARG);
    hipStream_t s = api_stream();
    MITOS_TRY(set_last_error(launch_hermitian<double>(reinterpret_cast<cx<double>*>(dg->fbox),
                                                      dg->N / 2 + 1, dg->N, 1, 0, s)));
    return set_last_error(api_sync());
}

// ---- misc ---------------------------------------------------------------------------
extern "C" void fft_wavevector(int x, int y, int z, int N, double delta_k, double* kx,
                               double* ky, double* kz, double* k) {
    wavevec(x, y, z, N, delta_k, *kx, *ky, *kz, *k);
}

extern "C" int mitos_last_error(void) { return g_last_error; }

extern "C" const char* mitos_strerror(int code) {
    switch (code) {
        case MITOS_OK: return "success";
        case MITOS_ERR_DIMS: return "non-matching grid dimensions";
        case MITOS_ERR_NOT_MOMENTUM: return "input field is not in momentum space";
        case MITOS_ERR_DEVICE: return "HIP device or runtime failure";
        case MITOS_ERR_UNSUPPORTED: return "size or kernel not supported on the device";
        case MITOS_ERR_ARG: return "invalid argument";
        case MITOS_ERR_ALLOC: return "device allocation failed";
        case MITOS_ERR_REAL_SPACE: return "input field is not in configuration space";
        case MITOS_ERR_TOO_LARGE:
            return "input larger than the fixed scratch of the allocation-free entry point";
    }
    return "unknown error";
}
