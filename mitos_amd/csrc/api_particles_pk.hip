// Drop-in particle steps (src/particle.c:87-129, src/mitos.c:827-863,
// analyse/src/mitos_read.c:278-376) and the power spectrum
// (src/calc_powerspec.c:32-217, include/calc_powerspec.h:26-36).
#include <math.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/mitos_hip.h"
#include "api_util.h"
#include "particles.h"
#include "powerspec.h"

using namespace mitos;

// Reference bin of calc_cross_powerspec for every integer n^2 in [1, 3(N/2)^2]:
// u is rounded to float and multiplied by (bins-1) in float
// (src/calc_powerspec.c:36-63).
extern "C" int mitos_pk_bin_table(int N, double boxlen, int bins, int* h) {
    if (N < 2 || bins < 1 || !h) return MITOS_ERR_ARG;
    const long n2max = 3L * (N / 2) * (N / 2);
    const double dk = 2 * M_PI / boxlen;
    const double max_k = sqrt(3) * dk * N / 2;
    const double min_k = dk;
    const double log_max_k = log(max_k);
    const double log_min_k = log(min_k);
    h[0] = -1;
    for (long n2 = 1; n2 <= n2max; n2++) {
        const double k = sqrt((double)n2) * dk;
        const float u = (log(k) - log_min_k) / (log_max_k - log_min_k);
        const float prod = (float)(bins - 1) * u;
        int bin = (int)floor(prod);
        if (bin < 0 || bin >= bins) bin = -1;
        h[n2] = bin;
    }
    return MITOS_OK;
}

namespace mitos {

const int* pk_bin_edges(int N, double boxlen, int bins, int* err) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, double, int>, int*> cache;
    int dev = 0;
    *err = MITOS_ERR_DEVICE;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto key = std::make_tuple(dev, N, boxlen, bins);
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    const long n2max = 3L * (N / 2) * (N / 2);
    std::vector<int> h(n2max + 1);
    *err = mitos_pk_bin_table(N, boxlen, bins, h.data());
    if (*err != MITOS_OK) return nullptr;
    // k = sqrt(n^2) dk grows by a relative step >= 1/(2 n^2) between integers,
    // far above the rounding of log, so the bin never decreases; checked here.
    std::vector<int> first(bins + 1, (int)(n2max + 1));
    int prev = 0;
    for (long n2 = 1; n2 <= n2max; n2++) {
        const int b = h[n2];
        if (b < prev || b < 0) {
            *err = MITOS_ERR_UNSUPPORTED;
            return nullptr;
        }
        for (int q = prev + 1; q <= b; q++) first[q] = (int)n2;
        if (n2 == 1)
            for (int q = 0; q <= b; q++) first[q] = 1;
        prev = b;
    }
    int* d = nullptr;
    *err = MITOS_ERR_ALLOC;
    if (hipMalloc(&d, sizeof(int) * first.size()) != hipSuccess) return nullptr;
    if (copy_sync(d, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    *err = MITOS_OK;
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = d;
    return d;
}

// multiplicity-weighted count of the half-grid modes with each integer n^2
static __global__ void k_n2_hist(int N, unsigned int* hist) {
    const int H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * N * H) return;
    const int z = (int)(i % H);
    const long row = i / H;
    const int nx = freq_index((int)(row / N), N), ny = freq_index((int)(row % N), N);
    const long n2 = (long)nx * nx + (long)ny * ny + (long)z * z;
    if (n2 == 0) return;
    atomicAdd(hist + n2, (z == 0 || z == N / 2) ? 1u : 2u);
}

int pk_geometry(int N, double boxlen, int bins, PkGeometry* out) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, double, int>, PkGeometry> cache;
    int dev = 0;
    MITOS_HIP_CHECK(hipGetDevice(&dev));
    auto key = std::make_tuple(dev, N, boxlen, bins);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *out = it->second;
        return MITOS_OK;
    }
    const long n2max = 3L * (N / 2) * (N / 2);
    std::vector<int> lut(n2max + 1);
    MITOS_TRY(mitos_pk_bin_table(N, boxlen, bins, lut.data()));
    unsigned int* dh = nullptr;
    MITOS_HIP_CHECK(hipMalloc(&dh, sizeof(unsigned int) * (n2max + 1)));
    MITOS_HIP_CHECK(hipMemset(dh, 0, sizeof(unsigned int) * (n2max + 1)));
    const long nmodes = (long)N * N * (N / 2 + 1);
    hipLaunchKernelGGL(k_n2_hist, dim3((unsigned)((nmodes + 255) / 256)), dim3(256), 0, 0, N, dh);
    std::vector<unsigned int> hist(n2max + 1);
    const hipError_t e = hipMemcpy(hist.data(), dh, sizeof(unsigned int) * (n2max + 1),
                                   hipMemcpyDeviceToHost);
    hipFree(dh);
    MITOS_HIP_CHECK(e);
    std::vector<double> ksum(bins, 0.0);
    std::vector<long long> obs(bins, 0);
    const double dk = 2 * M_PI / boxlen;
    for (long n2 = 1; n2 <= n2max; n2++) {
        if (!hist[n2] || lut[n2] < 0) continue;
        obs[lut[n2]] += hist[n2];
        ksum[lut[n2]] += (double)hist[n2] * (sqrt((double)n2) * dk);
    }
    double* dks = nullptr;
    long long* dob = nullptr;
    MITOS_HIP_CHECK(hipMalloc(&dks, sizeof(double) * bins));
    MITOS_HIP_CHECK(hipMalloc(&dob, sizeof(long long) * bins));
    MITOS_HIP_CHECK(copy_sync(dks, ksum.data(), sizeof(double) * bins, hipMemcpyHostToDevice));
    MITOS_HIP_CHECK(
        copy_sync(dob, obs.data(), sizeof(long long) * bins, hipMemcpyHostToDevice));
    PkGeometry g{dks, dob};
    cache[key] = g;
    *out = g;
    return MITOS_OK;
}

// src/calc_powerspec.c:84-89: averages, empty bins give 0/0 = NaN
void pk_finish(const double* sums, const long long* obs, int bins, double boxlen, double* k_out,
               double* p_out) {
    const double boxvol = boxlen * boxlen * boxlen;
    for (int i = 0; i < bins; i++) {
        const double o = (double)obs[i];
        k_out[i] = sums[2 * i + 1] / o;
        p_out[i] = sums[2 * i + 0] / o;
        p_out[i] /= boxvol;
    }
}

}  // namespace mitos

namespace {

template <typename T>
void cross_powerspec(int N, double boxlen, const void* box1, const void* box2, int bins,
                     double* k_in_bins, double* power_in_bins, long long* obs) {
    std::vector<double> sums(2 * bins, 0.0);
    int rc = MITOS_OK;
    if (!box1 || !box2 || bins <= 0 || !k_in_bins || !power_in_bins || !obs)
        rc = MITOS_ERR_ARG;
    else if (N < 2 || (N & 1))
        rc = MITOS_ERR_DIMS;
    if (rc == MITOS_OK)
        rc = pk_run<T, PkNoTransform>((const cx<T>*)box1, (const cx<T>*)box2, N, N / 2 + 1,
                                      boxlen, bins, PkNoTransform(), sums.data(), obs,
                                      api_stream(), nullptr, nullptr);
    set_last_error(rc);
    if (rc != MITOS_OK) {
        for (int i = 0; i < bins && k_in_bins && power_in_bins && obs; i++) {
            k_in_bins[i] = NAN;
            power_in_bins[i] = NAN;
            obs[i] = 0;
        }
        return;
    }
    pk_finish(sums.data(), obs, bins, boxlen, k_in_bins, power_in_bins);
}

inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" void calc_cross_powerspec(int N, double boxlen, const fftw_complex* box1,
                                     const fftw_complex* box2, int bins, double* k_in_bins,
                                     double* power_in_bins, int* obs_in_bins) {
    std::vector<long long> obs(bins > 0 ? bins : 1);
    cross_powerspec<double>(N, boxlen, box1, box2, bins, k_in_bins, power_in_bins, obs.data());
    for (int i = 0; i < bins && obs_in_bins; i++) obs_in_bins[i] = (int)obs[i];
}

extern "C" void calc_cross_powerspec_float(int N, double boxlen, const fftwf_complex* box1,
                                           const fftwf_complex* box2, int bins,
                                           double* k_in_bins, double* power_in_bins,
                                           int* obs_in_bins) {
    std::vector<long long> obs(bins > 0 ? bins : 1);
    cross_powerspec<float>(N, boxlen, box1, box2, bins, k_in_bins, power_in_bins, obs.data());
    for (int i = 0; i < bins && obs_in_bins; i++) obs_in_bins[i] = (int)obs[i];
}

extern "C" void calc_cross_powerspec_float_mpi(long int N, long int X0, long int NX,
                                               double boxlen, const fftwf_complex* box1,
                                               const fftwf_complex* box2, int bins,
                                               double* k_in_bins, double* power_in_bins,
                                               long int* obs_in_bins) {
    // one rank holds the whole grid: X0 = 0, NX = N
    if (X0 != 0 || NX != N) {
        set_last_error(MITOS_ERR_DIMS);
        for (int i = 0; i < bins; i++) {
            k_in_bins[i] = NAN;
            power_in_bins[i] = NAN;
            obs_in_bins[i] = 0;
        }
        return;
    }
    std::vector<long long> obs(bins > 0 ? bins : 1);
    cross_powerspec<float>((int)N, boxlen, box1, box2, bins, k_in_bins, power_in_bins,
                           obs.data());
    for (int i = 0; i < bins && obs_in_bins; i++) obs_in_bins[i] = (long int)obs[i];
}

// ---- Part 2: particles ---------------------------------------------------------------
extern "C" int mitos_lattice(int M, double boxlen, long long first_id, float* X, float* Y,
                             float* Z, long long* ids) {
    if (M < 1 || !X || !Y || !Z) return set_last_error(MITOS_ERR_ARG);
    const float len = (float)boxlen;
    const float spacing = len / M;  // src/particle.c:102-103
    const long long n = (long long)M * M * M;
    hipLaunchKernelGGL((k_lattice<float>), dim3(nblk(n)), dim3(256), 0, api_stream(), M, spacing,
                       first_id, X, Y, Z, ids);
    MITOS_TRY(set_last_error(launch_status()));
    return set_last_error(api_sync());
}

extern "C" int mitos_displace_tsc_dg(const struct distributed_grid* psi_x,
                                     const struct distributed_grid* psi_y,
                                     const struct distributed_grid* psi_z, long long npart,
                                     float* X, float* Y, float* Z) {
    if (!psi_x || !psi_y || !psi_z || !X || !Y || !Z || npart < 0)
        return set_last_error(MITOS_ERR_ARG);
    if (psi_x->N != psi_y->N || psi_y->N != psi_z->N) return set_last_error(MITOS_ERR_DIMS);
    if (psi_x->momentum_space || psi_y->momentum_space || psi_z->momentum_space)
        return set_last_error(MITOS_ERR_REAL_SPACE);
    const int N = psi_x->N;
    if (npart == 0) return MITOS_OK;
    // the gather indexes a grid with 32-bit element offsets
    if ((unsigned long long)(N + 2) * N * N > 0xffffffffULL)
        return set_last_error(MITOS_ERR_UNSUPPORTED);
    hipLaunchKernelGGL((k_displace<double>), dim3(nblk(npart)), dim3(256), 0, api_stream(),
                       psi_x->box, psi_y->box, psi_z->box, (unsigned)(N + 2), N, psi_x->boxlen,
                       npart, X, Y, Z);
    MITOS_TRY(set_last_error(launch_status()));
    return set_last_error(api_sync());
}

extern "C" int mitos_cic_overdensity(int N, double boxlen, long long npart, const float* X,
                                     const float* Y, const float* Z, double* box) {
    if (N < 2 || !box || npart <= 0 || !X || !Y || !Z) return set_last_error(MITOS_ERR_ARG);
    const long ncell = (long)N * N * N;
    unsigned long long* grid =
        (unsigned long long*)workspace(3, sizeof(unsigned long long) * ncell);
    if (!grid) return set_last_error(MITOS_ERR_ALLOC);
    hipStream_t s = api_stream();
    MITOS_HIP_CHECK(hipMemsetAsync(grid, 0, sizeof(unsigned long long) * ncell, s));
    hipLaunchKernelGGL(k_cic<double>, dim3(nblk(npart)), dim3(256), 0, s, X, Y, Z, npart, N, boxlen,
                       grid);
    MITOS_TRY(set_last_error(launch_status()));
    const double norm = ldexp(1.0, -MITOS_CIC_SHIFT) * ((double)ncell / (double)npart);
    hipLaunchKernelGGL((k_cic_to_delta<double>), dim3(nblk(ncell)), dim3(256), 0, s, grid, box,
                       ncell, (long)N, (long)N, norm);
    MITOS_TRY(set_last_error(launch_status()));
    return set_last_error(api_sync());
}
