// Whole-pipeline entry points: fused FP32/FP64 realization on one stream, and
// the reference's own call order through the drop-in functions (FP64 plain).
#include <math.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/mitos_hip.h"
#include "pipeline_kernels.h"
#include "powerspec.h"

using namespace mitos;

namespace mitos {


// ---- spectrum ------------------------------------------------------------------------
static double tophat_w(double x) {
    if (x < 1e-3) return 1.0 - x * x / 10.0;
    return 3.0 * (sin(x) - x * cos(x)) / (x * x * x);
}

// sigma_R^2 = int dk/(2 pi^2) k^2 P(k) W(kR)^2 with P = A k^n_s T_BBKS^2, Simpson in ln k
double sigma8_amplitude(double n_s, double gamma, double sigma8, double R) {
    const double lo = log(1e-6), hi = log(1e3);
    const int n = 200000;
    const double h = (hi - lo) / n;
    double sum = 0.0;
    for (int i = 0; i <= n; i++) {
        const double k = exp(lo + i * h);
        const double T = bbks_T(k, gamma);
        const double W = tophat_w(k * R);
        const double f = k * k * k * pow(k, n_s) * T * T * W * W / (2.0 * M_PI * M_PI);
        const double w = (i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
        sum += w * f;
    }
    return sigma8 * sigma8 / (sum * h / 3.0);
}

static double amplitude_cached(double n_s, double gamma, double sigma8) {
    static std::mutex mu;
    static std::map<std::tuple<double, double, double>, double> cache;
    auto key = std::make_tuple(n_s, gamma, sigma8);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const double A = sigma8_amplitude(n_s, gamma, sigma8, 8.0);
    cache[key] = A;
    return A;
}

// phi1(k) = -delta(k)/k^2 with delta = noise * sqrt(A k^n_s) T_BBKS(k)
struct AmpPhi1 {
    double A, n_s, gamma;
    __device__ __forceinline__ double operator()(double, double, double, double k) const {
#pragma clang fp contract(off)
        const double p = sqrt(A * pow(k, n_s));
        const double t = bbks_T(k, gamma);
        return (p * t) * (-1.0 / k / k);
    }
};

template <typename T>
__global__ void k_fill_u64(unsigned long long* p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// ---- launch helpers for the fused kernels ---------------------------------------------
template <typename T, int L, int AXIS, int DIR, int NIN, typename CODES, typename RAD>
struct MultiK;
template <typename T, int L, int AXIS, int DIR, int NIN, typename CODES, int... Rs>
struct MultiK<T, L, AXIS, DIR, NIN, CODES, Radices<Rs...>> {
    static int go(const MultiPtrs<T, NIN, CODES::n>& P, long hs, double dk, hipStream_t s) {
        constexpr int C = ColTile<T>::C;
        constexpr int E = LineCfg<L>::E;
        constexpr bool SPLIT = 2 * LdsMap<T, L, C, true>::SIZE * sizeof(T) > 65536;
        const int ncols = L / 2 + 1;
        const cx<T>* tw = twiddles<T>(L);
        if (!tw) return MITOS_ERR_ALLOC;
        hipLaunchKernelGGL((k_multi_strided<T, L, E, C, AXIS, DIR, SPLIT, NIN, CODES, Rs...>),
                           dim3((ncols + C - 1) / C, L), dim3(L / E * C), 0, s, P, hs, ncols, dk,
                           tw);
        return launch_status();
    }
};

template <typename T, int N, typename RAD>
struct HessZK;
template <typename T, int N, int... Rs>
struct HessZK<T, N, Radices<Rs...>> {
    static int go(const MultiPtrs<T, 6, 1>& P, long hs, double dk, T scale, hipStream_t s) {
        constexpr int L = N / 2;
        constexpr int E = LineCfg<L>::E;
        constexpr int TPL = L / E;
        constexpr int RW = RowsPerBlock<TPL>::value;
        const long nrows = (long)N * N;
        const cx<T>* tw = twiddles<T>(L);
        const cx<T>* twN = twiddles<T>(N);
        if (!tw || !twN) return MITOS_ERR_ALLOC;
        hipLaunchKernelGGL((k_hess_z<T, L, E, RW, Rs...>), dim3(nblk(nrows, RW)), dim3(TPL * RW),
                           0, s, P, nrows, hs, dk, scale, tw, twN);
        return launch_status();
    }
};

template <typename T, int N, typename RAD>
struct CombineK;
template <typename T, int N, int... Rs>
struct CombineK<T, N, Radices<Rs...>> {
    static int go(const cx<T>* D, const cx<T>* K, cx<T>* C0, cx<T>* C1, long hs,
                  const CombineParams& cp, hipStream_t s) {
        constexpr int L = N;
        constexpr int C = ColTile<T>::C;
        constexpr int E = LineCfg<L>::E;
        constexpr bool SPLIT = 2 * LdsMap<T, L, C, true>::SIZE * sizeof(T) > 65536;
        const int ncols = L / 2 + 1;
        const cx<T>* tw = twiddles<T>(L);
        if (!tw) return MITOS_ERR_ALLOC;
        hipLaunchKernelGGL((k_x_combine<T, L, E, C, SPLIT, Rs...>),
                           dim3((ncols + C - 1) / C, L), dim3(L / E * C), 0, s, D, K, C0, C1, hs,
                           ncols, cp, tw);
        return launch_status();
    }
};

// ---- the fused pipeline ------------------------------------------------------------------
template <typename T>
struct Pipe {
    int N = 0;
    long hs = 0;      // complex row stride (elements): N/2+1 rounded up to 128 bytes
    long csize = 0;   // elements of one complex grid
    cx<T>* K = nullptr;
    cx<T>* A = nullptr;  // 3 complex grids
    cx<T>* B = nullptr;  // 6 complex grids
    float *X = nullptr, *Y = nullptr, *Z = nullptr;
    double* pk_sums = nullptr;  // device [bins][2]
    long long* pk_obs = nullptr;
    int bins_alloc = 0;
    hipEvent_t ev[9] = {};
    double amplitude = 0;

    cx<T>* a(int i) { return A + (size_t)i * csize; }
    cx<T>* b(int i) { return B + (size_t)i * csize; }
    T* psi(int i) { return reinterpret_cast<T*>(b(3 + i)); }
    unsigned long long* cic() { return reinterpret_cast<unsigned long long*>(b(0)); }

    int create(int n) {
        N = n;
        const long H = N / 2 + 1;
        const long q = 128 / sizeof(cx<T>);
        hs = (H + q - 1) / q * q;
        csize = (long)N * N * hs;
        const long long np = (long long)N * N * N;
        MITOS_HIP_CHECK(hipMalloc(&K, sizeof(cx<T>) * csize));
        MITOS_HIP_CHECK(hipMalloc(&A, sizeof(cx<T>) * csize * 3));
        MITOS_HIP_CHECK(hipMalloc(&B, sizeof(cx<T>) * csize * 6));
        MITOS_HIP_CHECK(hipMalloc(&X, sizeof(float) * np));
        MITOS_HIP_CHECK(hipMalloc(&Y, sizeof(float) * np));
        MITOS_HIP_CHECK(hipMalloc(&Z, sizeof(float) * np));
        for (int i = 0; i < 9; i++) MITOS_HIP_CHECK(hipEventCreate(&ev[i]));
        return MITOS_OK;
    }

    void destroy() {
        hipDeviceSynchronize();
        hipFree(K);
        hipFree(A);
        hipFree(B);
        hipFree(X);
        hipFree(Y);
        hipFree(Z);
        hipFree(pk_sums);
        hipFree(pk_obs);
        for (int i = 0; i < 9; i++)
            if (ev[i]) hipEventDestroy(ev[i]);
    }

    int ensure_bins(int bins) {
        if (bins <= bins_alloc) return MITOS_OK;
        hipFree(pk_sums);
        hipFree(pk_obs);
        MITOS_HIP_CHECK(hipMalloc(&pk_sums, sizeof(double) * 2 * bins));
        MITOS_HIP_CHECK(hipMalloc(&pk_obs, sizeof(long long) * bins));
        bins_alloc = bins;
        return MITOS_OK;
    }

    int enqueue(const mitos_ic_config& cfg, hipStream_t s, bool timing) {
        const double L = cfg.boxlen;
        const double dk = 2 * M_PI / L;
        const double boxvol = L * L * L;
        const double sfwd = boxvol / ((long long)N * N * N);
        const long long np = (long long)N * N * N;
        const long H = N / 2 + 1;
        MITOS_TRY(ensure_bins(cfg.bins));
        amplitude = amplitude_cached(cfg.n_s, cfg.gamma, cfg.sigma8);
        auto mark = [&](int i) {
            if (timing) hipEventRecord(ev[i], s);
        };
        mark(0);
        // 1. noise -> phi1(k), Hermitian planes (src/grf.c, src/2lpt.c:75)
        Xo s0 = xo_seed(cfg.seed);
        NoiseJumps J;
        MITOS_TRY(noise_jumps(N, s0, &J));
        MITOS_TRY((launch_grf<T, AmpPhi1>(K, hs, N, L, 1, 0, J,
                                          AmpPhi1{amplitude, cfg.n_s, cfg.gamma}, s)));
        MITOS_TRY(launch_hermitian<T>(K, hs, N, 1, 0, s));
        mark(1);
        return with_size(N, [&]<int NN>() -> int {
            typedef typename LineCfg<NN>::R RX;
            typedef typename LineCfg<NN / 2>::R RZ;
            cx<T>* D = a(0);
            if (cfg.run2lpt) {
                // 2. x pass: A0 = X(phi1), A1 = X(i kx phi1), A2 = X(-kx^2 phi1)
                MultiPtrs<T, 1, 3> p2;
                p2.in[0] = K;
                p2.out[0] = a(0);
                p2.out[1] = a(1);
                p2.out[2] = a(2);
                MITOS_TRY((MultiK<T, NN, 0, 1, 1,
                                  OutCodes<MITOS_OUT(0, 0), MITOS_OUT(0, 1), MITOS_OUT(0, 2)>,
                                  RX>::go(p2, hs, dk, s)));
                // 3. y pass: six Hessian factors
                MultiPtrs<T, 3, 6> p3;
                p3.in[0] = a(0);
                p3.in[1] = a(1);
                p3.in[2] = a(2);
                for (int i = 0; i < 6; i++) p3.out[i] = b(i);
                MITOS_TRY((MultiK<T, NN, 1, 1, 3,
                                  OutCodes<MITOS_OUT(0, 0), MITOS_OUT(0, 1), MITOS_OUT(0, 2),
                                           MITOS_OUT(1, 0), MITOS_OUT(1, 1), MITOS_OUT(2, 0)>,
                                  RX>::go(p3, hs, dk, s)));
                mark(2);
                // 4. z pass: delta2 and its r2c row (src/2lpt.c:84-116, poisson.c:61)
                MultiPtrs<T, 6, 1> p4;
                for (int i = 0; i < 6; i++) p4.in[i] = b(i);
                p4.out[0] = D;
                MITOS_TRY((HessZK<T, NN, RZ>::go(p4, hs, dk, (T)(1.0 / boxvol), s)));
                // 5. y pass forward of delta2
                MITOS_TRY((strided_pass<T, NN, 1, -1>(D, D, hs, 1, 0, OpNone(), s)));
            } else {
                D = K;  // Zel'dovich: the second-order term is multiplied by zero
                mark(2);
            }
            mark(3);
            // 6. x pass: Poisson + combine + TSC window + first inverse pass
            CombineParams cp;
            cp.dk = dk;
            cp.sD = sfwd;
            cp.f1 = cfg.f1;
            cp.f2 = cfg.run2lpt ? cfg.f2 : 0.0;
            cp.wfac = 0.5 * L / N;
            cp.window = 3;
            MITOS_TRY((CombineK<T, NN, RX>::go(D, K, a(1), a(2), hs, cp, s)));
            // 7. y pass: E1 = Y(i ky C0), E2 = Y(C0), E0 = Y(C1)
            MultiPtrs<T, 2, 3> p7;
            p7.in[0] = a(1);
            p7.in[1] = a(2);
            p7.out[0] = b(1);
            p7.out[1] = b(2);
            p7.out[2] = b(0);
            MITOS_TRY((MultiK<T, NN, 1, 1, 2,
                              OutCodes<MITOS_OUT(0, 1), MITOS_OUT(0, 0), MITOS_OUT(1, 0)>,
                              RX>::go(p7, hs, dk, s)));
            // 8. z c2r: displacement grids (src/mitos.c:404-416)
            const T cs = (T)(1.0 / boxvol);
            MITOS_TRY((zc2r_pass<T, NN>(b(0), psi(0), hs, NN, 1, 0, 0, cs, OpNone(), s)));
            MITOS_TRY((zc2r_pass<T, NN>(b(1), psi(1), hs, NN, 1, 0, 0, cs, OpNone(), s)));
            MITOS_TRY((zc2r_pass<T, NN>(b(2), psi(2), hs, NN, 1, 0, 0, cs, OpIkz{dk}, s)));
            mark(4);
            // 9. lattice + sequential TSC displacement + wrap
            const float spacing = (float)L / NN;
            hipLaunchKernelGGL((k_displace<T, true>), dim3(nblk(np)), dim3(256), 0, s, psi(0),
                               psi(1), psi(2), (long)NN, NN, L, spacing, np, X, Y, Z);
            MITOS_TRY(launch_status());
            mark(5);
            // 10. CIC deposit (fixed point)
            hipLaunchKernelGGL((k_fill_u64<T>), dim3(nblk(np)), dim3(256), 0, s, cic(), (long)np);
            MITOS_TRY(launch_status());
            hipLaunchKernelGGL(k_cic, dim3(nblk(np)), dim3(256), 0, s, X, Y, Z, np, NN, L, cic());
            MITOS_TRY(launch_status());
            mark(6);
            // 11. density r2c (overdensity on load), 12. P(k) with CIC window
            CicLoader<T> ld{cic(), (long)NN, ldexp(1.0, -MITOS_CIC_SHIFT)};
            MITOS_TRY((zr2c_pass<T, NN>(ld, K, hs, 1, 0, OpNone(), s)));
            MITOS_TRY((strided_pass<T, NN, 1, -1>(K, K, hs, 1, 0, OpNone(), s)));
            MITOS_TRY((strided_pass<T, NN, 0, -1>(K, K, hs, 1, 0, OpNone(), s)));
            PkCicTransform xf{sfwd, 0.5 * L / NN};
            MITOS_TRY((pk_run<T, PkCicTransform>(K, K, NN, hs, L, cfg.bins, xf, nullptr, nullptr,
                                                 s, pk_sums, pk_obs)));
            mark(7);
            (void)H;
            return MITOS_OK;
        });
    }
};

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
    int rc;
    if (cfg->precision == MITOS_FP32) {
        c->pf = new Pipe<float>();
        rc = c->pf->create(cfg->N);
    } else {
        c->pd = new Pipe<double>();
        rc = c->pd->create(cfg->N);
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
    delete c;
    return MITOS_OK;
}

static int check_cfg(mitos_ic_context* c, const struct mitos_ic_config* cfg) {
    if (!c || !cfg) return MITOS_ERR_ARG;
    if (cfg->N != c->N || cfg->precision != c->precision) return MITOS_ERR_DIMS;
    if (cfg->bins < 1 || cfg->boxlen <= 0) return MITOS_ERR_ARG;
    return MITOS_OK;
}

extern "C" int mitos_ic_enqueue(mitos_ic_context* c, const struct mitos_ic_config* cfg,
                                void* stream) {
    MITOS_TRY(set_last_error(check_cfg(c, cfg)));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    c->bins = cfg->bins;
    c->boxlen = cfg->boxlen;
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

extern "C" int mitos_ic_positions(mitos_ic_context* c, float* X, float* Y, float* Z) {
    if (!c || !X || !Y || !Z) return set_last_error(MITOS_ERR_ARG);
    const size_t bytes = sizeof(float) * (size_t)c->N * c->N * c->N;
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

template <typename T>
__global__ void k_copy_psi(const T* src, long src_stride, T* dst, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const long g = i / n, j = i % n;
    dst[i] = src[g * src_stride + j];
}

extern "C" int mitos_ic_displacements(mitos_ic_context* c, void* psi) {
    if (!c || !psi) return set_last_error(MITOS_ERR_ARG);
    const long n = (long)c->N * c->N * c->N;
    if (c->pf)
        hipLaunchKernelGGL((k_copy_psi<float>), dim3(nblk(3 * n)), dim3(256), 0, c->stream,
                           c->pf->psi(0), 2 * c->pf->csize, (float*)psi, n);
    else
        hipLaunchKernelGGL((k_copy_psi<double>), dim3(nblk(3 * n)), dim3(256), 0, c->stream,
                           c->pd->psi(0), 2 * c->pd->csize, (double*)psi, n);
    MITOS_TRY(set_last_error(launch_status()));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MITOS_OK;
}

extern "C" int mitos_ic_density(mitos_ic_context* c, void* delta) {
    if (!c || !delta) return set_last_error(MITOS_ERR_ARG);
    const long N = c->N;
    const long n = N * N * N;
    const double norm = ldexp(1.0, -MITOS_CIC_SHIFT);
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
    const long long n = (long long)c->N * c->N * c->N;
    hipLaunchKernelGGL(k_ids, dim3(nblk(n)), dim3(256), 0, c->stream, n, 1LL, ids);
    MITOS_TRY(set_last_error(launch_status()));
    MITOS_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MITOS_OK;
}
