// The fused IC pipeline (struct Pipe) and its host-side tables, shared by the
// single-realization API (pipeline.hip) and the ensemble API (ensemble.hip).
#pragma once

#include <math.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/mitos_hip.h"
#include "pipeline_kernels.h"
#include "powerspec.h"

namespace mitos {


// ---- spectrum ------------------------------------------------------------------------
inline double tophat_w(double x) {
    if (x < 1e-3) return 1.0 - x * x / 10.0;
    return 3.0 * (sin(x) - x * cos(x)) / (x * x * x);
}

// sigma_R^2 = int dk/(2 pi^2) k^2 P(k) W(kR)^2 with P = A k^n_s T_BBKS^2, Simpson in ln k
inline double sigma8_amplitude(double n_s, double gamma, double sigma8, double R) {
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

// Host: device table w[i] = 1/sinc(0.5 k_i boxlen/N)^order for i in [0, N),
// k_i = freq_index(i)*2pi/boxlen.  The window of include/fft_kernels.h:198-219
// is separable, 1/W(k) = w[x] w[y] w[z] (k != 0).  Cached per (N, boxlen, order).
inline const double* window_table(int N, double boxlen, int order) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, double, int>, double*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto key = std::make_tuple(dev, N, boxlen, order);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<double> h(N);
    const double dk = 2 * M_PI / boxlen;
    for (int i = 0; i < N; i++) {
        const double k = freq_index(i, N) * dk;
        const double w = sinc_d(0.5 * k * boxlen / N);
        h[i] = 1.0 / pow(w, order);
    }
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * N) != hipSuccess) return nullptr;
    if (copy_sync(d, h.data(), sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    cache[key] = d;
    return d;
}

// Device table S[n] = 3/4 + cos(2 pi n/N)/4, n in [0, N): the DFT of the
// periodic lattice stencil (1/8, 3/4, 1/8).  Cached per (device, N, type).
template <typename V>
inline const V* smooth_table(int N) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, V*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto key = std::make_pair(dev, N);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<V> h(N);
    for (int n = 0; n < N; n++) {
        // exact values at the quarter points
        double c;
        if (4 * n == N || 4 * n == 3 * N) c = 0.0;
        else if (2 * n == N) c = -1.0;
        else c = cos(2.0 * M_PI * n / N);
        h[n] = (V)(0.75 + 0.25 * c);
    }
    V* d = nullptr;
    if (hipMalloc(&d, sizeof(V) * N) != hipSuccess) return nullptr;
    if (copy_sync(d, h.data(), sizeof(V) * N, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    cache[key] = d;
    return d;
}

// True when every lattice coordinate q*((float)L/N) (float, src/particle.c:
// 102-109) sits exactly on grid node q in the reference's cell units
// (double x*N/boxlen, src/grids_interp.c:269-271).  Then the TSC window of an
// unmoved coordinate is exactly (1/8, 3/4, 1/8) and the lattice stencils
// reproduce gridTSC_dg; otherwise (e.g. 2000 Mpc/h over 1536 cells, a spacing
// float rounds) the moves gather 27 taps at the actual positions.
inline bool lattice_on_nodes(int N, double L) {
    const float spacing = (float)L / N;
    for (int q = 0; q < N; q++) {
        const float x = (float)q * spacing;
        if ((double)x * N / L != (double)q) return false;
    }
    return true;
}

inline double amplitude_cached(double n_s, double gamma, double sigma8) {
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

// Device table over the integer n^2 = |k|^2/dk^2 of the amplitude taking white
// noise to phi1(k) = -delta(k)/k^2, delta = noise * sqrt(A k^n_s) T_BBKS(k)
// (the primordial kernel of src/mitos.c:224, the transfer kernel of
// src/generate_grids.c:52 and kernel_inv_poisson of src/2lpt.c:75).
inline std::vector<double> phi1_host(int N, double boxlen, double A, double n_s, double gamma) {
    const long n2max = 3L * (N / 2) * (N / 2);
    std::vector<double> h(n2max + 1);
    const double dk = 2 * M_PI / boxlen;
    h[0] = 0.0;
    for (long n2 = 1; n2 <= n2max; n2++) {
        const double k = sqrt((double)n2) * dk;
        const double p = sqrt(A * pow(k, n_s));
        const double t = bbks_T(k, gamma);
        h[n2] = (p * t) * (-1.0 / k / k);
    }
    return h;
}

inline const double* phi1_table(int N, double boxlen, double A, double n_s, double gamma) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, double, double, double, double>, double*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto key = std::make_tuple(dev, N, boxlen, A, n_s, gamma);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (cache.size() > 64) {  // bounded: one entry per (size, spectrum)
        hipDeviceSynchronize();
        for (auto& kv : cache) hipFree(kv.second);
        cache.clear();
    }
    const std::vector<double> h = phi1_host(N, boxlen, A, n_s, gamma);
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return nullptr;
    if (copy_sync(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    cache[key] = d;
    return d;
}

// The table of phi1_table times the noise factor sqrt(L^3/2) of k_grf (the
// same double expression), rounded to float: the FP32 noise kernel's values
inline const float* phi1_table_f(int N, double boxlen, double A, double n_s, double gamma) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, double, double, double, double>, float*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto key = std::make_tuple(dev, N, boxlen, A, n_s, gamma);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (cache.size() > 64) {
        hipDeviceSynchronize();
        for (auto& kv : cache) hipFree(kv.second);
        cache.clear();
    }
    const std::vector<double> h = phi1_host(N, boxlen, A, n_s, gamma);
    const double boxvol = boxlen * boxlen * boxlen;
    const double factor = sqrt(boxvol / 2);
    std::vector<float> f(h.size());
    for (size_t i = 0; i < h.size(); i++) f[i] = (float)(factor * h[i]);
    float* d = nullptr;
    if (hipMalloc(&d, sizeof(float) * f.size()) != hipSuccess) return nullptr;
    if (copy_sync(d, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    cache[key] = d;
    return d;
}

// K = noise * sqrt(L^3/2) * amp[n^2] (the scaling of src/grf.c:33 and the
// phi1 table), DC = 0; noise dense N x N x (N/2+1)
template <typename T>
__global__ void k_noise_in(const cx<T>* __restrict__ noise, cx<T>* K, long hs, int N,
                           double boxlen, const double* __restrict__ amp) {
    const int H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * N * H) return;
    const int z = (int)(i % H);
    const long row = i / H;
    const int x = (int)(row / N), y = (int)(row % N);
    const int nx = freq_index(x, N), ny = freq_index(y, N);
    const long n2 = (long)nx * nx + (long)ny * ny + (long)z * z;
    const double f = sqrt(boxlen * boxlen * boxlen / 2) * amp[n2];
    const cx<T> v = noise[i];
    K[row * hs + z] = n2 ? mk<T>((T)((double)v.re * f), (T)((double)v.im * f)) : mk<T>(0, 0);
}

template <typename T>
__global__ void k_fill_u64(unsigned long long* p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// ---- launch helpers for the fused kernels ---------------------------------------------
// Line configuration and tile width of the register-heavy multi-transform
// kernels: 8 (or 12) values per thread and 64-byte tile rows, so that the
// input tile, the transformed copy and the radix temporaries stay in VGPRs.
// Multi-output passes on float lines of 1024: 16 values per thread, so a
// line is one wave (fewer exchange steps), and 128-byte tile rows (whole
// lines; work-groups of 1024 threads, one per CU); the combine pass keeps the
// smaller line configuration (MULTI = false).
template <typename T, int L, bool MULTI = true>
struct HeavyCfg {
    static constexpr bool WIDE = MULTI && sizeof(T) == 4 && L == 1024;
    typedef typename std::conditional<WIDE, LineCfg<L>, LineCfgS<L>>::type CFG;
    static constexpr int E = CFG::E;
    static constexpr int TPL = L / E;
    static constexpr int C0 = (WIDE ? 128 : 64) / (2 * (int)sizeof(T));
    static constexpr int C = (TPL * C0 <= 1024) ? C0 : 1024 / TPL;
};
template <typename T, int L, int AXIS, int DIR, int NIN, typename CODES,
          typename RAD = typename HeavyCfg<T, L>::CFG::R>
struct MultiK;
template <typename T, int L, int AXIS, int DIR, int NIN, typename CODES, int... Rs>
struct MultiK<T, L, AXIS, DIR, NIN, CODES, Radices<Rs...>> {
    static int go(const MultiPtrs<T, NIN, CODES::n>& P, long hs, double dk, hipStream_t s) {
        constexpr int C = HeavyCfg<T, L>::C;
        constexpr int E = HeavyCfg<T, L>::E;
        // one exchange buffer per component unless that leaves less than two
        // work-groups per CU; the wide form runs one per CU by its registers
        // anyway (1024 threads at about 124), so it takes both (139 KB)
        constexpr bool SPLIT = !HeavyCfg<T, L>::WIDE &&
                               2 * LdsMap<T, L, C, true>::SIZE * sizeof(T) > 81920;
        const int ncols = L / 2 + 1;
        const cx<T>* tw = twiddles<T>(L);
        if (!tw) return MITOS_ERR_ALLOC;
        constexpr int KO = (AXIS == 0 && L >= 64) ? 2 : 1;
        hipLaunchKernelGGL((k_multi_strided<T, L, E, C, AXIS, DIR, SPLIT, NIN, CODES, KO, Rs...>),
                           dim3((ncols + C - 1) / C, L / KO), dim3(L / E * C), 0, s, P, hs, ncols,
                           dk, tw);
        return launch_status();
    }
};

template <typename T, int N, typename RAD = typename LineCfgS<N / 2>::R>
struct HessZK;
template <typename T, int N, int... Rs>
struct HessZK<T, N, Radices<Rs...>> {
    // delta2 rows written as reals (row stride rs) for a separate z pass
    static int go(const MultiPtrs<T, 6, 1>& P, long hs, double dk, T scale, T* real, long rs,
                  hipStream_t s) {
        constexpr int L = N / 2;
        constexpr int E = LineCfgS<L>::E;
        constexpr int TPL = L / E;
        // a row of a wave or more: one row per work-group (no barrier waits
        // on other rows' waves)
        constexpr int RW = TPL >= 64 ? 1 : RowsPerBlock<TPL>::value;
        const long nrows = (long)N * N;
        const cx<T>* tw = twiddles<T>(L);
        const cx<T>* twN = twiddles<T>(N);
        if (!tw || !twN) return MITOS_ERR_ALLOC;
        hipLaunchKernelGGL((k_hess_z<T, L, E, RW, Rs...>), dim3(nblk(nrows, RW)),
                           dim3(TPL * RW), 0, s, P, nrows, hs, dk, scale, tw, twN, real, rs);
        return launch_status();
    }
};

template <typename T, int N, typename RAD = typename HeavyCfg<T, N, false>::CFG::R>
struct CombineK;
template <typename T, int N, int... Rs>
struct CombineK<T, N, Radices<Rs...>> {
    static int go(const cx<T>* D, const cx<T>* K, cx<T>* C0, cx<T>* C1, long hs,
                  const CombineParams& cp, hipStream_t s) {
        constexpr int L = N;
        constexpr int C = HeavyCfg<T, L, false>::C;
        constexpr int E = HeavyCfg<T, L, false>::E;
        constexpr bool SPLIT = 2 * LdsMap<T, L, C, true>::SIZE * sizeof(T) > 81920;
        const int ncols = L / 2 + 1;
        const cx<T>* tw = twiddles<T>(L);
        if (!tw) return MITOS_ERR_ALLOC;
        hipLaunchKernelGGL((k_x_combine<T, L, E, C, SPLIT, Rs...>),
                           dim3((ncols + C - 1) / C, L), dim3(L / E * C), 0, s, D, K, C0, C1, hs,
                           ncols, cp, tw);
        return launch_status();
    }
};

// ---- FFT schedule autotuner ------------------------------------------------------------
// The plain passes of the pipeline have kFftVariants schedules each (radix
// split, tile width, rows per work-group; see fft_launch.h).  The first
// pipeline of a (device, N, precision) times every variant on its own grids
// and keeps the fastest; later pipelines of the same key reuse the choice,
// so every run in the process uses the same kernels (ensembles and unbatched
// runs give identical bits).  MITOS_FFT_SCHEDULE=default skips the timing.
enum TuneKind { TUNE_Y_FWD = 0, TUNE_X_FWD = 1, TUNE_Z_C2R = 2, TUNE_Z_R2C = 3, TUNE_KINDS = 4 };

struct FftSchedule {
    int v[TUNE_KINDS] = {0, 0, 0, 0};
    float ms[TUNE_KINDS][kFftVariants] = {};
};

inline std::mutex& schedule_mutex() {
    static std::mutex mu;
    return mu;
}
inline std::map<std::tuple<int, int, int>, FftSchedule>& schedule_table() {
    static std::map<std::tuple<int, int, int>, FftSchedule> t;
    return t;
}

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
    long long pcap = 0;   // particles X, Y, Z hold
    long long npart = 0;  // particles of the last realization (M^3)
    double* pk_sums = nullptr;  // device [bins][2]
    long long* pk_obs = nullptr;
    double* pk_part = nullptr;  // per-work-group P(k) partials of this pipeline
    long long* pk_cnt = nullptr;
    int bins_alloc = 0;
    hipEvent_t ev[9] = {};
    double amplitude = 0;
    // optional per-kernel profile of the last run: name, algorithmic bytes,
    // start/stop events on the run's stream
    bool prof = false;
    int nprof = 0;
    const char* pname[32];
    double pbytes[32];
    hipEvent_t pev[33];
    template <typename F>
    int timed(hipStream_t s, const char* name, double bytes, F&& f) {
        if (prof && nprof < 32) {
            hipEventRecord(pev[nprof], s);
            const int rc = f();
            pname[nprof] = name;
            pbytes[nprof] = bytes;
            nprof++;
            hipEventRecord(pev[nprof], s);
            return rc;
        }
        return f();
    }

    FftSchedule sched;
    bool psi_stencilled = false;  // the displacement grids hold the stencil form
    // workspace of the two-pass lattice deposit (float weights, N >= 128);
    // null: the deposit adds its windows with atomics
    void* cic_ws = nullptr;
    // the adjoint's fixed-point scatter accumulators (two int64 grids and two
    // maxima), allocated at the first backward run
    unsigned long long* adj_ws = nullptr;
    // the FP32 displacement stages its taps in LDS (MITOS_TSC_DIRECT=1: reads
    // them from the grids, the same bits; for tests)
    bool tsc_staged = true;
    // the FP32 displacement over 8 x 8 x 16 tiles (MITOS_TSC_KERNEL=wave: the
    // one-wave row kernel)
    bool tsc_tile = true;

    // time every schedule variant of the plain passes (or reuse the choice
    // made for this device, N and precision)
    int autotune(hipStream_t s) {
        int dev = 0;
        MITOS_HIP_CHECK(hipGetDevice(&dev));
        const auto key = std::make_tuple(dev, N, (int)sizeof(T));
        std::lock_guard<std::mutex> lock(schedule_mutex());
        auto& table = schedule_table();
        auto it = table.find(key);
        if (it != table.end()) {
            sched = it->second;
            return MITOS_OK;
        }
        FftSchedule sc;
        const char* env = getenv("MITOS_FFT_SCHEDULE");
        if (!(env && strcmp(env, "default") == 0)) {
            hipEvent_t e0, e1;
            MITOS_HIP_CHECK(hipEventCreate(&e0));
            MITOS_HIP_CHECK(hipEventCreate(&e1));
            int rc = with_size(N, [&]<int NN>() -> int {
                auto time_it = [&](auto&& launch, float* ms) -> int {
                    MITOS_TRY(launch());  // warm-up
                    MITOS_HIP_CHECK(hipEventRecord(e0, s));
                    for (int r = 0; r < 3; r++) MITOS_TRY(launch());
                    MITOS_HIP_CHECK(hipEventRecord(e1, s));
                    MITOS_HIP_CHECK(hipEventSynchronize(e1));
                    MITOS_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
                    *ms /= 3;
                    return MITOS_OK;
                };
                CicLoader<T> ld{cic(), (long)NN, 1.0};
                for (int v = 0; v < kFftVariants; v++) {
                    // out of place, as the runs use them
                    MITOS_TRY(time_it([&] {
                        return strided_pass_tuned<T, NN, 1, -1>(v, K, a(0), hs, 1, 0, OpNone(),
                                                                s);
                    }, &sc.ms[TUNE_Y_FWD][v]));
                    MITOS_TRY(time_it([&] {
                        return strided_pass_tuned<T, NN, 0, -1>(v, a(0), K, hs, 1, 0, OpNone(),
                                                                s);
                    }, &sc.ms[TUNE_X_FWD][v]));
                    MITOS_TRY(time_it([&] {
                        // the displacement grids' layout of the default form
                        constexpr int HO = sizeof(T) == 4 ? MITOS_PSI_SO : MITOS_PSI_HO;
                        return zc2r_pass_tuned<T, NN, HO>(v, b(0), psi(0), hs, prs(), 1, 0, 0,
                                                          (T)1, OpNone(), s);
                    }, &sc.ms[TUNE_Z_C2R][v]));
                    MITOS_TRY(time_it([&] {
                        return zr2c_pass_tuned<T, NN>(v, ld, K, hs, 1, 0, OpNone(), s);
                    }, &sc.ms[TUNE_Z_R2C][v]));
                }
                return MITOS_OK;
            });
            hipEventDestroy(e0);
            hipEventDestroy(e1);
            MITOS_TRY(rc);
            for (int k = 0; k < TUNE_KINDS; k++)
                for (int v = 1; v < kFftVariants; v++)
                    if (sc.ms[k][v] < sc.ms[k][sc.v[k]]) sc.v[k] = v;
        }
        table[key] = sc;
        sched = sc;
        return MITOS_OK;
    }

    // Mean time of a plain 3D r2c (real rows of prs() -> K) and c2r
    // (K -> real) over reps runs each, with the tuned schedules; uses the
    // pipeline's grids as scratch.
    int fft_timing(hipStream_t s, int reps, float* r2c_ms, float* c2r_ms) {
        hipEvent_t e0, e1, e2;
        MITOS_HIP_CHECK(hipEventCreate(&e0));
        MITOS_HIP_CHECK(hipEventCreate(&e1));
        MITOS_HIP_CHECK(hipEventCreate(&e2));
        T* real = psi(0);
        const long rs = prs();
        const int rc = with_size(N, [&]<int NN>() -> int {
            RealRowLoader<T> ld{real, rs, 0};
            auto r2c = [&]() -> int {
                MITOS_TRY((zr2c_pass_tuned<T, NN>(sched.v[TUNE_Z_R2C], ld, K, hs, 1, 0, OpNone(),
                                                  s)));
                // out of place (an in-place strided pass streams slower)
                MITOS_TRY((strided_pass_tuned<T, NN, 1, -1>(sched.v[TUNE_Y_FWD], K, a(0), hs, 1,
                                                            0, OpNone(), s)));
                return strided_pass_tuned<T, NN, 0, -1>(sched.v[TUNE_X_FWD], a(0), K, hs, 1, 0,
                                                        OpNone(), s);
            };
            auto c2r = [&]() -> int {
                MITOS_TRY((strided_pass_tuned<T, NN, 0, 1>(sched.v[TUNE_X_FWD], K, a(0), hs, 1, 0,
                                                           OpNone(), s)));
                MITOS_TRY((strided_pass_tuned<T, NN, 1, 1>(sched.v[TUNE_Y_FWD], a(0), a(1), hs, 1,
                                                           0, OpNone(), s)));
                return zc2r_pass_tuned<T, NN, 0>(sched.v[TUNE_Z_C2R], a(1), real, hs, rs, 1, 0, 0,
                                                 (T)1, OpNone(), s);
            };
            MITOS_TRY(r2c());
            MITOS_TRY(c2r());
            MITOS_HIP_CHECK(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) MITOS_TRY(r2c());
            MITOS_HIP_CHECK(hipEventRecord(e1, s));
            for (int r = 0; r < reps; r++) MITOS_TRY(c2r());
            MITOS_HIP_CHECK(hipEventRecord(e2, s));
            MITOS_HIP_CHECK(hipEventSynchronize(e2));
            MITOS_HIP_CHECK(hipEventElapsedTime(r2c_ms, e0, e1));
            MITOS_HIP_CHECK(hipEventElapsedTime(c2r_ms, e1, e2));
            *r2c_ms /= reps;
            *c2r_ms /= reps;
            return MITOS_OK;
        });
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        hipEventDestroy(e2);
        return rc;
    }

    cx<T>* a(int i) { return A + (size_t)i * csize; }
    cx<T>* b(int i) { return B + (size_t)i * csize; }
    T* psi(int i) { return reinterpret_cast<T*>(b(3 + i)); }
    // Row stride (reals) of the displacement and delta2 grids: one complex row,
    // so every real row starts on a 128-byte boundary as the complex rows do.
    long prs() const { return 2 * hs; }
    unsigned long long* cic() { return reinterpret_cast<unsigned long long*>(b(0)); }
    // fixed-point unit of the lattice deposit (float weights in the FP32 pipeline)
    static int cic_shift() { return sizeof(T) == 4 ? MITOS_CIC_SHIFT_F32 : MITOS_CIC_SHIFT; }
    // overdensity scale of the fixed-point grid: 2^-shift * N^3 / npart
    double cic_norm() const {
        const double n3 = (double)N * N * N;
        const double f = (npart == 0 || (double)npart == n3) ? 1.0 : n3 / (double)npart;
        return ldexp(1.0, -cic_shift()) * f;
    }

    // npart_cap: the most lattice particles a realization may have (0: N^3)
    int create(int n, long long npart_cap = 0) {
        N = n;
        const long H = N / 2 + 1;
        const long q = 128 / sizeof(cx<T>);
        hs = (H + q - 1) / q * q;
        csize = (long)N * N * hs;
        // a displacement grid (rows of prs() = 2 hs reals, halo included) fits
        // in a complex grid
        if (N + MITOS_PSI_HO + 1 > 2 * hs) return MITOS_ERR_UNSUPPORTED;
        const long long ng = (long long)N * N * N;
        pcap = npart_cap > ng ? npart_cap : ng;
        MITOS_HIP_CHECK(hipMalloc(&K, sizeof(cx<T>) * csize));
        MITOS_HIP_CHECK(hipMalloc(&A, sizeof(cx<T>) * csize * 3));
        MITOS_HIP_CHECK(hipMalloc(&B, sizeof(cx<T>) * csize * 6));
        MITOS_HIP_CHECK(hipMalloc(&X, sizeof(float) * pcap));
        MITOS_HIP_CHECK(hipMalloc(&Y, sizeof(float) * pcap));
        MITOS_HIP_CHECK(hipMalloc(&Z, sizeof(float) * pcap));
        typedef typename std::conditional<sizeof(T) == 4, float, double>::type M;
        const size_t wsb = cic_stream_bytes<M>(N);
        const char* direct_env = getenv("MITOS_TSC_DIRECT");
        tsc_staged = !(direct_env && direct_env[0] == '1');
        const char* kern_env = getenv("MITOS_TSC_KERNEL");
        tsc_tile = !(kern_env && strcmp(kern_env, "wave") == 0);
        const char* atomic_env = getenv("MITOS_CIC_ATOMIC");
        const bool force_atomic = atomic_env && atomic_env[0] == '1';
        if (wsb && !force_atomic && hipMalloc(&cic_ws, wsb) != hipSuccess) {
            (void)hipGetLastError();  // no room: the atomic deposit needs none
            cic_ws = nullptr;
        }
        for (int i = 0; i < 9; i++) MITOS_HIP_CHECK(hipEventCreate(&ev[i]));
        for (int i = 0; i < 33; i++) MITOS_HIP_CHECK(hipEventCreate(&pev[i]));
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
        hipFree(pk_part);
        hipFree(pk_cnt);
        hipFree(cic_ws);
        hipFree(adj_ws);
        for (int i = 0; i < 9; i++)
            if (ev[i]) hipEventDestroy(ev[i]);
        for (int i = 0; i < 33; i++)
            if (pev[i]) hipEventDestroy(pev[i]);
    }

    int ensure_bins(int bins) {
        if (bins <= bins_alloc) return MITOS_OK;
        hipFree(pk_sums);
        hipFree(pk_obs);
        hipFree(pk_part);
        hipFree(pk_cnt);
        const int nb = pk_blocks((long)N * N * (N / 2 + 1));
        MITOS_HIP_CHECK(hipMalloc(&pk_sums, sizeof(double) * 2 * bins));
        MITOS_HIP_CHECK(hipMalloc(&pk_obs, sizeof(long long) * bins));
        MITOS_HIP_CHECK(hipMalloc(&pk_part, sizeof(double) * 4 * bins * nb));
        MITOS_HIP_CHECK(hipMalloc(&pk_cnt, sizeof(long long) * bins * nb));
        bins_alloc = bins;
        return MITOS_OK;
    }

    // One realization on stream s.  amp_tab: device phi1 amplitude table to use
    // instead of the cached one; pk_dst/obs_dst: device destinations of the
    // P(k) sums and counts instead of pk_sums/pk_obs.  No allocation and no
    // host synchronisation happen here once the tables of cfg exist, so the
    // call can be captured into a graph.
    // noise_in: unit-variance complex white noise (dense N x N x (N/2+1) of cx<T>)
    // used instead of the seed's stream.  stencil: -1 as the precision
    // (FP32 applies the lattice TSC stencils in k-space, FP64 gathers 27 taps),
    // 1 forces the stencil form (the form the adjoint differentiates).
    int enqueue(const mitos_ic_config& cfg, hipStream_t s, bool timing,
                const double* amp_tab = nullptr, double* pk_dst = nullptr,
                long long* obs_dst = nullptr, const cx<T>* noise_in = nullptr,
                int stencil = -1) {
        // particle lattice M^3 (src/particle.c:87-129); the lattice-tile
        // kernels (stencil form, tiled moves and deposits) need M = N
        const int Ml = cfg.M > 0 ? cfg.M : N;
        const long long nl = (long long)Ml * Ml * Ml;
        if (nl > pcap) return MITOS_ERR_DIMS;
        if (Ml != N && (noise_in || stencil > 0)) return MITOS_ERR_UNSUPPORTED;
        npart = nl;
        // the stencil form is exact only for particles on grid nodes
        const bool STENCIL =
            Ml == N && (stencil < 0 ? (sizeof(T) == 4 && lattice_on_nodes(N, cfg.boxlen))
                                    : stencil != 0);
        psi_stencilled = STENCIL;
        const double L = cfg.boxlen;
        const double dk = 2 * M_PI / L;
        const double boxvol = L * L * L;
        const double sfwd = boxvol / ((long long)N * N * N);
        const long long np = (long long)N * N * N;
        // algorithmic bytes: one half-complex grid, one real grid
        const double G = (double)N * N * (N / 2 + 1) * sizeof(cx<T>);
        const double Rg = (double)np * sizeof(T);
        MITOS_TRY(ensure_bins(cfg.bins));
        amplitude = amplitude_cached(cfg.n_s, cfg.gamma, cfg.sigma8);
        nprof = 0;
        auto mark = [&](int i) {
            if (timing) hipEventRecord(ev[i], s);
        };
        mark(0);
        // 1. noise -> phi1(k), Hermitian planes (src/grf.c, src/2lpt.c:75)
        Xo s0 = xo_seed(cfg.seed);
        NoiseJumps J;
        MITOS_TRY(noise_jumps(N, s0, &J));
        const double* amp = amp_tab ? amp_tab : phi1_table(N, L, amplitude, cfg.n_s, cfg.gamma);
        if (!amp) return MITOS_ERR_ALLOC;
        // FP64 keeps correctly rounded Box-Muller; FP32 rounds every value to float
        MITOS_TRY(timed(s, "grf_noise", G, [&] {
            if (noise_in) {
                hipLaunchKernelGGL((k_noise_in<T>), dim3(nblk((long)N * N * (N / 2 + 1))),
                                   dim3(256), 0, s, noise_in, K, hs, N, L, amp);
                return launch_status();
            }
            if constexpr (sizeof(T) == 4) {
                if (!amp_tab) {
                    // an unbatched FP32 run: the prescaled float table (the
                    // same values as the double table's product)
                    const float* ampf = phi1_table_f(N, L, amplitude, cfg.n_s, cfg.gamma);
                    if (!ampf) return MITOS_ERR_ALLOC;
                    return launch_grf<T, AmpTableF, false>(K, hs, N, L, 1, 0, J, AmpTableF{ampf},
                                                           s);
                }
            }
            return launch_grf<T, AmpTable, sizeof(T) == 8>(K, hs, N, L, 1, 0, J, AmpTable{amp}, s);
        }));
        MITOS_TRY(timed(s, "hermitian_planes", 0.0,
                        [&] { return launch_hermitian<T>(K, hs, N, 1, 0, s); }));
        mark(1);
        return with_size(N, [&]<int NN>() -> int {
            cx<T>* D = a(0);
            if (cfg.run2lpt) {
                // 2. x pass: A0 = X(phi1), A1 = X(i kx phi1), A2 = X(-kx^2 phi1)
                MultiPtrs<T, 1, 3> p2;
                p2.in[0] = K;
                p2.out[0] = a(0);
                p2.out[1] = a(1);
                p2.out[2] = a(2);
                MITOS_TRY(timed(s, "hessian_x_1to3", 4 * G, [&] {
                    return MultiK<T, NN, 0, 1, 1,
                                  OutCodes<MITOS_OUT(0, 0), MITOS_OUT(0, 1), MITOS_OUT(0, 2)>>::go(
                        p2, hs, dk, s);
                }));
                // 3. y pass: six Hessian factors
                MultiPtrs<T, 3, 6> p3;
                p3.in[0] = a(0);
                p3.in[1] = a(1);
                p3.in[2] = a(2);
                for (int i = 0; i < 6; i++) p3.out[i] = b(i);
                MITOS_TRY(timed(s, "hessian_y_3to6", 9 * G, [&] {
                    return MultiK<T, NN, 1, 1, 3,
                                  OutCodes<MITOS_OUT(0, 0), MITOS_OUT(0, 1), MITOS_OUT(0, 2),
                                           MITOS_OUT(1, 0), MITOS_OUT(1, 1), MITOS_OUT(2, 0)>>::go(
                        p3, hs, dk, s);
                }));
                mark(2);
                // 4. z pass: delta2 and its r2c row (src/2lpt.c:84-116, poisson.c:61)
                MultiPtrs<T, 6, 1> p4;
                for (int i = 0; i < 6; i++) p4.in[i] = b(i);
                p4.out[0] = D;
                // delta2(x) rows into a(1) as reals, then their r2c into D
                T* d2 = reinterpret_cast<T*>(a(1));
                const long rs2 = prs();
                MITOS_TRY(timed(s, "hessian_z_delta2", 6 * G + Rg, [&] {
                    return HessZK<T, NN>::go(p4, hs, dk, (T)(1.0 / boxvol), d2, rs2, s);
                }));
                MITOS_TRY(timed(s, "delta2_z_r2c", Rg + G, [&] {
                    RealRowLoader<T> ld{d2, rs2, 0};
                    return zr2c_pass_tuned<T, NN>(sched.v[TUNE_Z_R2C], ld, D, hs, 1, 0, OpNone(),
                                                  s);
                }));
                // 5. y pass forward of delta2, out of place into a(2)
                MITOS_TRY(timed(s, "delta2_y_fwd", 2 * G, [&] {
                    return strided_pass_tuned<T, NN, 1, -1>(sched.v[TUNE_Y_FWD], D, a(2), hs, 1,
                                                            0, OpNone(), s);
                }));
                D = a(2);
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
            cp.win = window_table(NN, L, 3);
            if (!cp.win) return MITOS_ERR_ALLOC;
            // the stencil form applies the lattice TSC stencils in k-space
            // (k_displace_lattice_s)
            cp.smooth = nullptr;
            if (STENCIL) {
                cp.smooth = smooth_table<double>(NN);
                if (!cp.smooth) return MITOS_ERR_ALLOC;
            }
            // C0, C1 into a(0), a(1) (D is a(2) or K)
            MITOS_TRY(timed(s, "combine_x", 4 * G, [&] {
                return CombineK<T, NN>::go(D, K, a(0), a(1), hs, cp, s);
            }));
            // 7. y pass: E1 = Y(i ky C0) (FP32: i ky S(ky) C0), E2 = Y(C0), E0 = Y(C1)
            MultiPtrs<T, 2, 3> p7;
            p7.in[0] = a(0);
            p7.in[1] = a(1);
            p7.out[0] = b(1);
            p7.out[1] = b(2);
            p7.out[2] = b(0);
            if (STENCIL) {
                p7.smooth = smooth_table<T>(NN);
                if (!p7.smooth) return MITOS_ERR_ALLOC;
            }
            MITOS_TRY(timed(s, "displacement_y_2to3", 5 * G, [&] {
                if (STENCIL)
                    return MultiK<T, NN, 1, 1, 2,
                                  OutCodes<MITOS_OUT(0, 3), MITOS_OUT(0, 0),
                                           MITOS_OUT(1, 0)>>::go(p7, hs, dk, s);
                return MultiK<T, NN, 1, 1, 2,
                              OutCodes<MITOS_OUT(0, 1), MITOS_OUT(0, 0), MITOS_OUT(1, 0)>>::go(
                    p7, hs, dk, s);
            }));
            // 8. z c2r: displacement grids (src/mitos.c:404-416); the
            // stencil form without the z halo (its moves read no z taps)
            const T cs = (T)(1.0 / boxvol);
            auto psi_c2r = [&](const cx<T>* in, T* out, auto op) -> int {
                if (STENCIL)
                    return zc2r_pass_tuned<T, NN, MITOS_PSI_SO>(sched.v[TUNE_Z_C2R], in, out, hs,
                                                                prs(), 1, 0, 0, cs, op, s);
                return zc2r_pass_tuned<T, NN, MITOS_PSI_HO>(sched.v[TUNE_Z_C2R], in, out, hs,
                                                            prs(), 1, 0, 0, cs, op, s);
            };
            MITOS_TRY(timed(s, "displacement_z_c2r", G + Rg,
                            [&] { return psi_c2r(b(0), psi(0), OpNone()); }));
            MITOS_TRY(timed(s, "displacement_z_c2r", G + Rg,
                            [&] { return psi_c2r(b(1), psi(1), OpNone()); }));
            MITOS_TRY(timed(s, "displacement_z_c2r", G + Rg,
                            [&] { return psi_c2r(b(2), psi(2), OpIkz{dk}); }));
            mark(4);
            // 9-10. lattice + sequential TSC displacement + wrap + CIC deposit
            const float spacing = (float)L / Ml;  // src/particle.c:102-103
            // the two-pass deposit of the lattice particles stores every
            // grid cell, the one-pass and the adjoint-mode deposits add to a
            // cleared grid (far-particle lists use K, free after combine_x)
            typedef typename std::conditional<sizeof(T) == 4, float, double>::type M;
            const bool two_pass = !noise_in && cic_ws && NN % 32 == 0 && Ml == NN &&
                                  cic_list_bytes<M>(NN) <= sizeof(cx<T>) * (size_t)csize;
            if (!two_pass)
                MITOS_TRY(timed(s, "cic_clear", 8.0 * np, [&] {
                    hipLaunchKernelGGL((k_fill_u64<T>), dim3(nblk(np)), dim3(256), 0, s, cic(),
                                       (long)np);
                    return launch_status();
                }));
            mark(5);
            // FP32 moves the particles in float arithmetic, FP64 in double
            if (noise_in) {
                // the forward of the adjoint: positions stay in T up to the deposit
                MITOS_TRY(timed(s, "move_deposit", 3 * Rg + 12.0 * np + 8.0 * np, [&] {
                    hipLaunchKernelGGL((k_move_deposit_s<T>), dim3(nblk(np)), dim3(256), 0, s,
                                       psi(0), psi(1), psi(2), (unsigned)(prs()),
                                       NN, L, spacing, X, Y, Z, cic(), cic_shift());
                    return launch_status();
                }));
            } else {
                // reads 3 displacement grids, writes 3 float coordinates
                MITOS_TRY(timed(s, "displace_tsc", 3 * Rg + 12.0 * nl, [&] {
                    constexpr int NT = NN >= 256 ? 256 : 64;
                    const int nch = (NN + NT - 1) / NT;
                    const dim3 grid((unsigned)((long)NN * NN * nch));
                    const unsigned rs = (unsigned)(prs());
                    if (Ml != NN) {
                        // an M^3 lattice off the grid nodes: the 27-tap moves
                        // at each particle's position
                        const int nchm = (Ml + NT - 1) / NT;
                        hipLaunchKernelGGL((k_displace_lattice<T, M, NT>),
                                           dim3((unsigned)((long)Ml * Ml * nchm)), dim3(NT), 0, s,
                                           psi(0), psi(1), psi(2), rs, NN, Ml, nchm, L, spacing,
                                           X, Y, Z);
                    } else if (STENCIL && NN % 16 == 0 && tsc_tile) {
                        auto tile = [&](auto kern, int txn, int ty, int tz, int pp) {
                            const dim3 gridt(
                                (unsigned)((long)(NN / txn) * (NN / ty) * (NN / tz)));
                            hipLaunchKernelGGL(kern, gridt, dim3(txn / pp * ty * tz), 0, s,
                                               psi(0), psi(1), psi(2), rs, NN, (T)L, spacing, L,
                                               X, Y, Z, tsc_staged ? 1 : 0);
                        };
                        // 8 x 8 x 16 tiles, four particles per thread; windows of at
                        // most 40 x rows and 576 rows of 16 z (a limit of 608 left
                        // 0.2% of the tiles of a 1.95 Mpc/h cell at z = 0 amplitude
                        // to the direct reads), four work-groups per CU
                        tile(k_displace_tile<T, 8, 8, 16, 4, 40, 576, 12, 0>, 8, 8, 16, 4);
                    } else if (STENCIL && NN % 64 == 0) {
                        constexpr int TX = 16, PPT = 8;
                        const dim3 gridw((unsigned)((long)NN * NN / PPT * (NN / 64)));
                        hipLaunchKernelGGL((k_displace_lattice_w<T, TX, PPT, 16, 6>), gridw,
                                           dim3(64), 0, s, psi(0), psi(1), psi(2), rs, NN,
                                           (T)L, spacing, L, X, Y, Z, tsc_staged ? 1 : 0);
                    } else if (STENCIL) {
                        constexpr int TX = NN % 16 == 0 ? 16 : 8;
                        constexpr int PPT = 4;
                        const dim3 gridp((unsigned)((long)NN * NN / PPT * nch));
                        hipLaunchKernelGGL((k_displace_lattice_s<T, NT, TX, PPT>), gridp,
                                           dim3(NT), 0, s, psi(0), psi(1), psi(2), rs, NN, nch,
                                           (T)L, spacing, L, X, Y, Z);
                    } else if (sizeof(T) == 8 && NN % 16 == 0 && tsc_tile) {
                      if constexpr (sizeof(T) == 8) {
                        // the reference's 27-tap moves over 8 x 8 x 16 tiles with the
                        // taps staged in LDS (same positions as the per-particle
                        // kernel, which MITOS_TSC_KERNEL=wave selects)
                        auto tile = [&](auto kern, int txn, int ty, int tz, int pp) {
                            const dim3 gridt(
                                (unsigned)((long)(NN / txn) * (NN / ty) * (NN / tz)));
                            hipLaunchKernelGGL(kern, gridt, dim3(txn / pp * ty * tz), 0, s,
                                               psi(0), psi(1), psi(2), rs, NN, spacing, L, X,
                                               Y, Z);
                        };
                        // 256 window rows (37 KB of LDS) and eight rows in flight per
                        // 32 lanes; four work-groups per CU (128 registers)
                        tile(k_displace_tile_ref<8, 8, 16, 4, 256, 8>, 8, 8, 16, 4);
                      }
                    } else {
                        hipLaunchKernelGGL((k_displace_lattice<T, M, NT>), grid, dim3(NT), 0, s,
                                           psi(0), psi(1), psi(2), rs, NN, NN, nch, L, spacing,
                                           X, Y, Z);
                    }
                    return launch_status();
                }));
                // reads 3 float coordinates, writes the grid (two passes: the
                // windows through the workspace, then summed into the grid)
                MITOS_TRY(timed(s, "cic_deposit", two_pass ? 12.0 * np : 12.0 * nl + 8.0 * np, [&] {
                    if constexpr (sizeof(M) == 4) {
                        if (two_pass) {
                            MITOS_TRY((cic_lattice_launch<M, true>(X, Y, Z, NN, L, cic(), cic_ws,
                                                                   K, s)));
                            return launch_status();
                        }
                    }
                    if (NN % 32 == 0 && Ml == NN) {
                        MITOS_TRY((cic_lattice_launch<M, false>(X, Y, Z, NN, L, cic(), nullptr,
                                                                nullptr, s)));
                        return launch_status();
                    }
                    hipLaunchKernelGGL(k_cic<M>, dim3(nblk(nl)), dim3(256), 0, s, X, Y, Z, nl, NN, L,
                                       cic());
                    return launch_status();
                }));
                if constexpr (sizeof(M) == 4) {
                    if (two_pass) {
                        MITOS_TRY(timed(s, "cic_gather", 8.0 * np, [&] {
                            MITOS_TRY(cic_gather_launch<M>(NN, cic(), cic_ws, s));
                            return launch_status();
                        }));
                        MITOS_TRY(timed(s, "cic_far", 0.0, [&] {
                            MITOS_TRY(cic_far_launch<M>(X, Y, Z, NN, L, cic(), cic_ws, K, s));
                            return launch_status();
                        }));
                    }
                }
            }
            mark(6);
            // 11. density r2c (overdensity on load), 12. P(k) with CIC window
            // delta = sum_w * N^3 / npart - 1 (analyse/src/mitos_read.c:363-376)
            CicLoader<T> ld{cic(), (long)NN, cic_norm()};
            MITOS_TRY(timed(s, "density_z_r2c", 8.0 * np + G, [&] {
                return zr2c_pass_tuned<T, NN>(sched.v[TUNE_Z_R2C], ld, K, hs, 1, 0, OpNone(), s);
            }));
            // y and x out of place through a(0) (free once the displacement
            // grids are made)
            MITOS_TRY(timed(s, "density_y_fwd", 2 * G, [&] {
                return strided_pass_tuned<T, NN, 1, -1>(sched.v[TUNE_Y_FWD], K, a(0), hs, 1, 0,
                                                        OpNone(), s);
            }));
            MITOS_TRY(timed(s, "density_x_fwd", 2 * G, [&] {
                return strided_pass_tuned<T, NN, 0, -1>(sched.v[TUNE_X_FWD], a(0), K, hs, 1, 0,
                                                        OpNone(), s);
            }));
            const double* w2 = window_table(NN, L, 2);
            if (!w2) return MITOS_ERR_ALLOC;
            PkCicTransform xf{sfwd, w2, NN};
            MITOS_TRY(timed(s, "pk_bins", G, [&] {
                return pk_run<T, PkCicTransform>(K, K, NN, hs, L, cfg.bins, xf, nullptr, nullptr,
                                                 s, pk_dst ? pk_dst : pk_sums,
                                                 obs_dst ? obs_dst : pk_obs, pk_part, pk_cnt);
            }));
            mark(7);
            return MITOS_OK;
        });
    }
};

}  // namespace mitos
