// The CIC overdensity of mitos_cic_overdensity without any allocation, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  The fixed-point accumulation grid lives in a
// file-scope device array of fixed size (MITOS_NOALLOC_MAX_CELLS cells, held
// by every GPU the library runs on until the process ends); its device
// address is looked up once per device.  Calls that overlap on one GPU share
// that array and race.
#include <mutex>

#include "../../include/mitos_hip.h"
#include "api_util.h"
#include "particles.h"

using namespace mitos;

namespace {

constexpr int kMaxDevices = 64;

// fixed-point CIC grid (2^-40 units) of at most MITOS_NOALLOC_MAX_CELLS cells
__device__ __align__(256) unsigned long long scratch[MITOS_NOALLOC_MAX_CELLS];

std::mutex g_mu;
unsigned long long* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once)
int scratch_ptr(unsigned long long** out) {
    int dev = 0;
    MITOS_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return MITOS_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        MITOS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned long long*>(p);
    }
    *out = g_scratch[dev];
    return MITOS_OK;
}

__global__ void k_zero_u64(unsigned long long* p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" int mitos_cic_overdensity_noalloc(int N, double boxlen, long long npart,
                                             const float* X, const float* Y, const float* Z,
                                             double* box, void* stream) {
    if (N < 2 || !box || npart <= 0 || !X || !Y || !Z) return set_last_error(MITOS_ERR_ARG);
    const long ncell = (long)N * N * N;
    if (ncell > (long)MITOS_NOALLOC_MAX_CELLS) return set_last_error(MITOS_ERR_TOO_LARGE);
    unsigned long long* grid = nullptr;
    MITOS_TRY(set_last_error(scratch_ptr(&grid)));
    hipStream_t s = (hipStream_t)stream;  // NULL: the null stream, then a wait
    // the array keeps the previous call's sums: clear it (a fill kernel, so
    // the call can be captured into a graph)
    hipLaunchKernelGGL(k_zero_u64, dim3(nblk(ncell)), dim3(256), 0, s, grid, ncell);
    MITOS_TRY(set_last_error(launch_status()));
    hipLaunchKernelGGL(k_cic<double>, dim3(nblk(npart)), dim3(256), 0, s, X, Y, Z, npart, N,
                       boxlen, grid);
    MITOS_TRY(set_last_error(launch_status()));
    const double norm = ldexp(1.0, -MITOS_CIC_SHIFT) * ((double)ncell / (double)npart);
    hipLaunchKernelGGL((k_cic_to_delta<double>), dim3(nblk(ncell)), dim3(256), 0, s, grid, box,
                       ncell, (long)N, (long)N, norm);
    MITOS_TRY(set_last_error(launch_status()));
    if (!stream) MITOS_HIP_CHECK(hipStreamSynchronize(nullptr));
    return MITOS_OK;
}
