// Whole-pipeline entry points: fused FP32/FP64 realization on one stream, and
// the reference's own call order through the drop-in functions (FP64 plain).
#include "pipe.h"

using namespace mitos;

namespace mitos {

template struct Pipe<float>;
template struct Pipe<double>;

}  // namespace mitos

// ---- C API ------------------------------------------------------------------------------
struct mitos_ic_context {
    int N;
    int precision;
    int bins;
    double boxlen;
    hipStream_t stream;
    hipStream_t last = nullptr;  // stream of the last enqueued realization
    hipEvent_t order = nullptr;  // orders a new realization after the last one
    bool own_stream;
    bool timing;
    Pipe<float>* pf;
    Pipe<double>* pd;
    double amplitude() const { return pf ? pf->amplitude : pd->amplitude; }
};

extern "C" double mitos_sigma8_amplitude(double n_s, double gamma, double sigma8, double R) {
    return sigma8_amplitude(n_s, gamma, sigma8, R);
}

extern "C" int mitos_ic_create(const struct mitos_ic_config* cfg, mitos_ic_context** out) {
    if (!cfg || !out) return set_last_error(MITOS_ERR_ARG);
    if (!size_supported(cfg->N)) return set_last_error(MITOS_ERR_UNSUPPORTED);
    if (cfg->precision != MITOS_FP32 && cfg->precision != MITOS_FP64)
        return set_last_error(MITOS_ERR_ARG);
    if (cfg->M < 0 || cfg->M > 2048) return set_last_error(MITOS_ERR_ARG);
    pool_trim();  // cached drop-in blocks give way to the pipeline's grids
    mitos_ic_context* c = new mitos_ic_context();
    c->N = cfg->N;
    c->precision = cfg->precision;
    c->bins = cfg->bins;
    c->boxlen = cfg->boxlen;
    c->own_stream = true;
    c->timing = true;
    c->pf = nullptr;
    c->pd = nullptr;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return set_last_error(MITOS_ERR_DEVICE);
    }
    if (hipEventCreateWithFlags(&c->order, hipEventDisableTiming) != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return set_last_error(MITOS_ERR_DEVICE);
    }
    // the particle arrays hold the config's M^3 lattice (at least N^3)
    const long long Ml = cfg->M > 0 ? cfg->M : cfg->N;
    int rc;
    if (cfg->precision == MITOS_FP32) {
        c->pf = new Pipe<float>();
        rc = c->pf->create(cfg->N, Ml * Ml * Ml);
        if (rc == MITOS_OK) rc = c->pf->autotune(c->stream);
    } else {
        c->pd = new Pipe<double>();
        rc = c->pd->create(cfg->N, Ml * Ml * Ml);
        if (rc == MITOS_OK) rc = c->pd->autotune(c->stream);
    }
    if (rc != MITOS_OK) {
        mitos_ic_destroy(c);
        return set_last_error(rc == MITOS_ERR_DEVICE ? MITOS_ERR_ALLOC : rc);
    }
    *out = c;
    return MITOS_OK;
}

extern "C" int mitos_ic_destroy(mitos_ic_context* c) {
    if (!c) return MITOS_ERR_ARG;
    if (c->pf) {
        c->pf->destroy();
        delete c->pf;
    }
    if (c->pd) {
        c->pd->destroy();
        delete c->pd;
    }
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    if (c->order) hipEventDestroy(c->order);
    delete c;
    return MITOS_OK;
}

// a realization enqueued on a caller stream has finished before the
// context's own stream reads the pipeline buffers
static int wait_caller_stream(mitos_ic_context* c) {
    if (c->last && c->last != c->stream) MITOS_HIP_CHECK(hipStreamSynchronize(c->last));
    return MITOS_OK;
}

// Work about to be issued on stream s reuses the pipeline buffers (K, A, B,
// X/Y/Z, P(k) sums) of the last realization, which may still run on another
// stream: s waits for it on the device (an event, so the order also holds
// inside a graph capture of s), and s becomes the last stream.
static int order_after_last(mitos_ic_context* c, hipStream_t s) {
    if (c->last && c->last != s) {
        MITOS_HIP_CHECK(hipEventRecord(c->order, c->last));
        MITOS_HIP_CHECK(hipStreamWaitEvent(s, c->order, 0));
    }
    c->last = s;
    return MITOS_OK;
}

static int check_cfg(mitos_ic_context* c, const struct mitos_ic_config* cfg) {
    if (!c || !cfg) return MITOS_ERR_ARG;
    if (cfg->N != c->N || cfg->precision != c->precision) return MITOS_ERR_DIMS;
    if (cfg->bins < 1 || cfg->boxlen <= 0 || cfg->M < 0 || cfg->M > 2048) return MITOS_ERR_ARG;
    return MITOS_OK;
}

// lattice particles of the last realization (N^3 before the first)
static long long ctx_npart(const mitos_ic_context* c) {
    const long long n = c->pf ? c->pf->npart : c->pd->npart;
    return n > 0 ? n : (long long)c->N * c->N * c->N;
}

namespace mitos {
Pipe<float>* ctx_pipe_f(mitos_ic_context* c) { return c->pf; }
Pipe<double>* ctx_pipe_d(mitos_ic_context* c) { return c->pd; }
hipStream_t ctx_stream(mitos_ic_context* c) { return c->stream; }
int ctx_order(mitos_ic_context* c, hipStream_t s) { return order_after_last(c, s); }
int ctx_use_cfg(mitos_ic_context* c, const mitos_ic_config* cfg) {
    MITOS_TRY(check_cfg(c, cfg));
    c->bins = cfg->bins;
    c->boxlen = cfg->boxlen;
    return MITOS_OK;
}
}  // namespace mitos

extern "C" int mitos_ic_enqueue(mitos_ic_context* c, const struct mitos_ic_config* cfg,
                                void* stream) {
    MITOS_TRY(set_last_error(check_cfg(c, cfg)));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    c->bins = cfg->bins;
    c->boxlen = cfg->boxlen;
    MITOS_TRY(set_last_error(order_after_last(c, s)));
    int rc = c->pf ? c->pf->enqueue(*cfg, s, c->timing && !stream)
                   : c->pd->enqueue(*cfg, s, c->timing && !stream);
    return set_last_error(rc);
}

extern "C" int mitos_ic_fetch_pk(mitos_ic_context* c, struct mitos_ic_result* res) {
    if (!c || !res) return set_last_error(MITOS_ERR_ARG);
    const int bins = c->bins;
    std::vector<double> sums(2 * bins);
    std::vector<long long> obs(bins);
    double* ds = c->pf ? c->pf->pk_sums : c->pd->pk_sums;
    long long* dob = c->pf ? c->pf->pk_obs : c->pd->pk_obs;
    MITOS_TRY(set_last_error(wait_caller_stream(c)));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    MITOS_HIP_CHECK(hipMemcpy(sums.data(), ds, sizeof(double) * 2 * bins, hipMemcpyDeviceToHost));
    MITOS_HIP_CHECK(hipMemcpy(obs.data(), dob, sizeof(long long) * bins, hipMemcpyDeviceToHost));
    std::vector<double> kk(bins), pp(bins);
    pk_finish(sums.data(), obs.data(), bins, c->boxlen, kk.data(), pp.data());
    if (res->k_in_bins) memcpy(res->k_in_bins, kk.data(), sizeof(double) * bins);
    if (res->power_in_bins) memcpy(res->power_in_bins, pp.data(), sizeof(double) * bins);
    if (res->obs_in_bins) memcpy(res->obs_in_bins, obs.data(), sizeof(long long) * bins);
    res->amplitude = c->amplitude();
    return MITOS_OK;
}

extern "C" int mitos_ic_run(mitos_ic_context* c, const struct mitos_ic_config* cfg,
                            struct mitos_ic_result* res) {
    MITOS_TRY(mitos_ic_enqueue(c, cfg, nullptr));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    MITOS_HIP_CHECK(hipGetLastError());
    if (res) {
        MITOS_TRY(mitos_ic_fetch_pk(c, res));
        hipEvent_t* ev = c->pf ? c->pf->ev : c->pd->ev;
        for (int i = 0; i < 7; i++) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            res->stage_ms[i] = ms;
        }
        float tot = 0;
        hipEventElapsedTime(&tot, ev[0], ev[7]);
        res->stage_ms[7] = tot;
    }
    return MITOS_OK;
}

extern "C" int mitos_ic_schedule(mitos_ic_context* c, int kind, float* ms) {
    if (!c || kind < 0 || kind >= TUNE_KINDS) return -1;
    const FftSchedule& sc = c->pf ? c->pf->sched : c->pd->sched;
    if (ms)
        for (int v = 0; v < kFftVariants; v++) ms[v] = sc.ms[kind][v];
    return sc.v[kind];
}

extern "C" int mitos_ic_fft_timing(mitos_ic_context* c, int reps, float* r2c_ms,
                                   float* c2r_ms) {
    if (!c || reps < 1 || !r2c_ms || !c2r_ms) return set_last_error(MITOS_ERR_ARG);
    MITOS_TRY(set_last_error(order_after_last(c, c->stream)));
    const int rc = c->pf ? c->pf->fft_timing(c->stream, reps, r2c_ms, c2r_ms)
                         : c->pd->fft_timing(c->stream, reps, r2c_ms, c2r_ms);
    return set_last_error(rc);
}

extern "C" int mitos_set_fft_schedule(int N, int precision, const int* variants) {
    if (!variants || !size_supported(N) || (precision != MITOS_FP32 && precision != MITOS_FP64))
        return set_last_error(MITOS_ERR_ARG);
    FftSchedule sc;
    for (int k = 0; k < TUNE_KINDS; k++) {
        if (variants[k] < 0 || variants[k] >= kFftVariants) return set_last_error(MITOS_ERR_ARG);
        sc.v[k] = variants[k];
    }
    int dev = 0;
    MITOS_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(schedule_mutex());
    schedule_table()[std::make_tuple(dev, N, precision == MITOS_FP32 ? 4 : 8)] = sc;
    return MITOS_OK;
}

extern "C" int mitos_ic_profile(mitos_ic_context* c, int enable) {
    if (!c) return set_last_error(MITOS_ERR_ARG);
    if (c->pf) c->pf->prof = enable != 0;
    if (c->pd) c->pd->prof = enable != 0;
    return MITOS_OK;
}

extern "C" int mitos_ic_profile_entry(mitos_ic_context* c, int i, const char** name, float* ms,
                                      double* bytes) {
    if (!c) return -1;
    const int n = c->pf ? c->pf->nprof : c->pd->nprof;
    if (i < 0 || i >= n) return n;
    hipEvent_t* ev = c->pf ? c->pf->pev : c->pd->pev;
    MITOS_HIP_CHECK(hipEventSynchronize(ev[i + 1]));
    float t = 0;
    MITOS_HIP_CHECK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
    if (name) *name = c->pf ? c->pf->pname[i] : c->pd->pname[i];
    if (ms) *ms = t;
    if (bytes) *bytes = c->pf ? c->pf->pbytes[i] : c->pd->pbytes[i];
    return n;
}

extern "C" int mitos_ic_positions(mitos_ic_context* c, float* X, float* Y, float* Z) {
    if (!c || !X || !Y || !Z) return set_last_error(MITOS_ERR_ARG);
    MITOS_TRY(set_last_error(wait_caller_stream(c)));
    const size_t bytes = sizeof(float) * (size_t)ctx_npart(c);
    float* src[3] = {c->pf ? c->pf->X : c->pd->X, c->pf ? c->pf->Y : c->pd->Y,
                     c->pf ? c->pf->Z : c->pd->Z};
    float* dst[3] = {X, Y, Z};
    for (int i = 0; i < 3; i++)
        MITOS_HIP_CHECK(hipMemcpyAsync(dst[i], src[i], bytes, hipMemcpyDeviceToDevice, c->stream));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MITOS_OK;
}

__global__ void k_ids(long long n, long long first_id, long long* ids) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) ids[p] = p + first_id;  // src/particle.c:111 with FirstID = 1
}

// k-space operator dividing out the lattice TSC stencils that the FP32
// pipeline folds into its displacement grids (bit 0: S(kx), 1: S(ky), 2:
// S(kz)), with the inverse-FFT normalisation
struct OpUnstencil {
    const double* S;
    int mask;
    double norm;
    template <typename T>
    __device__ __forceinline__ cx<T> pre(cx<T> v, int x, int y, int z) const {
        double f = norm;
        if (mask & 1) f /= S[x];
        if (mask & 2) f /= S[y];
        if (mask & 4) f /= S[z];
        return mk<T>((T)(v.re * f), (T)(v.im * f));
    }
    template <typename T>
    __device__ __forceinline__ cx<T> post(cx<T> v, int, int, int) const {
        return v;
    }
};

// dense (3, N, N, N) copy of the halo-padded displacement grids
template <typename T>
__global__ void k_copy_psi(const T* src, long src_stride, long rs, T* dst, long n, int N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const long g = i / n, j = i % n;
    const long row = j / N, z = j % N;
    dst[i] = src[g * src_stride + row * rs + MITOS_PSI_HO + z];
}

// FP32: the grids hold S_xyz psi_x, S_yz psi_y, S_z psi_z (see
// k_displace_lattice_s); psi is recovered by r2c, division by the stencils and
// c2r, in the pipeline's scratch grids (K and A, free after a run)
template <typename T>
static int unstencil_psi(Pipe<T>* P, int N, T* out, hipStream_t s) {
    const double* S = smooth_table<double>(N);
    if (!S) return MITOS_ERR_ALLOC;
    const GridLayout gin{N, P->hs, P->prs()};
    const GridLayout gout{N, P->hs, (long)N};
    const double norm = 1.0 / ((double)N * N * N);
    const int masks[3] = {7, 6, 4};
    for (int i = 0; i < 3; i++) {
        RealRowLoader<T> ld{P->psi(i) + MITOS_PSI_SO, gin.rs, 0};
        MITOS_TRY((r2c_3d<T>(ld, P->K, gin, 1, 0, OpNone(), OpNone(), OpNone(), s)));
        MITOS_TRY((c2r_3d<T>(P->K, P->a(0), out + (size_t)i * N * N * N, gout, 1, 0, 0, (T)1,
                             OpUnstencil{S, masks[i], norm}, OpNone(), OpNone(), s)));
    }
    return MITOS_OK;
}

extern "C" int mitos_ic_displacements(mitos_ic_context* c, void* psi) {
    if (!c || !psi) return set_last_error(MITOS_ERR_ARG);
    MITOS_TRY(set_last_error(wait_caller_stream(c)));
    const long n = (long)c->N * c->N * c->N;
    if (c->pf && c->pf->psi_stencilled)
        MITOS_TRY(set_last_error(unstencil_psi(c->pf, c->N, (float*)psi, c->stream)));
    else if (c->pd && c->pd->psi_stencilled)
        MITOS_TRY(set_last_error(unstencil_psi(c->pd, c->N, (double*)psi, c->stream)));
    else if (c->pf)
        hipLaunchKernelGGL((k_copy_psi<float>), dim3(nblk(3 * n)), dim3(256), 0, c->stream,
                           c->pf->psi(0), 2 * c->pf->csize, c->pf->prs(), (float*)psi, n, c->N);
    else
        hipLaunchKernelGGL((k_copy_psi<double>), dim3(nblk(3 * n)), dim3(256), 0, c->stream,
                           c->pd->psi(0), 2 * c->pd->csize, c->pd->prs(), (double*)psi, n, c->N);
    MITOS_TRY(set_last_error(launch_status()));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MITOS_OK;
}

extern "C" int mitos_ic_density(mitos_ic_context* c, void* delta) {
    if (!c || !delta) return set_last_error(MITOS_ERR_ARG);
    MITOS_TRY(set_last_error(wait_caller_stream(c)));
    const long N = c->N;
    const long n = N * N * N;
    const double norm = c->pf ? c->pf->cic_norm() : c->pd->cic_norm();
    if (c->pf)
        hipLaunchKernelGGL((k_cic_to_delta<float>), dim3(nblk(n)), dim3(256), 0, c->stream,
                           c->pf->cic(), (float*)delta, n, N, N, norm);
    else
        hipLaunchKernelGGL((k_cic_to_delta<double>), dim3(nblk(n)), dim3(256), 0, c->stream,
                           c->pd->cic(), (double*)delta, n, N, N, norm);
    MITOS_TRY(set_last_error(launch_status()));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MITOS_OK;
}

extern "C" int mitos_ic_ids(mitos_ic_context* c, long long* ids) {
    if (!c || !ids) return set_last_error(MITOS_ERR_ARG);
    const long long n = ctx_npart(c);
    hipLaunchKernelGGL(k_ids, dim3(nblk(n)), dim3(256), 0, c->stream, n, 1LL, ids);
    MITOS_TRY(set_last_error(launch_status()));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MITOS_OK;
}
