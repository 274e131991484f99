// Binned cross power spectrum (src/calc_powerspec.c:32-90) with a fixed,
// thread-order-independent reduction: every wavefront folds equal-bin values
// with a butterfly sum, lane 0 adds them into its own compensated (Neumaier)
// accumulators in LDS, work-groups combine their wavefronts in order, and a
// second kernel folds the work-group partials in order.
#pragma once

#include "fft_launch.h"
#include "grf.h"
#include "runtime.h"

namespace mitos {

__device__ __forceinline__ void kadd(double& s, double& c, double v) {
#pragma clang fp contract(off)
    const double t = s + v;
    if (fabs(s) >= fabs(v))
        c += (s - t) + v;
    else
        c += (v - t) + s;
    s = t;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ long long wave_sum_i(long long v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// value transform of the binned modes: mode (x, y, z) is multiplied by
// row(x, y) * col(z) (a separable real factor)
struct PkNoTransform {
    __device__ __forceinline__ double row(int, int) const { return 1.0; }
    __device__ __forceinline__ double col(int) const { return 1.0; }
};

// Bin of a mode with integer n^2 = nx^2 + ny^2 + nz^2 > 0: the reference's
// bin (src/calc_powerspec.c:58-59) is nondecreasing in n^2, so it is the
// largest b with first[b] <= n^2.  A float estimate of the log-spaced bin is
// corrected against the table (usually no step).
__device__ __forceinline__ int pk_bin_of(long n2, const int* first, int bins, float cinv) {
    int b = (int)(__logf((float)n2) * cinv);
    b = b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
    while (b < bins - 1 && n2 >= first[b + 1]) b++;
    while (b > 0 && n2 < first[b]) b--;
    return b;
}

// Partials layout: part[(blk*bins + b)*2 + {sum, compensation}] of the power.
// Mode counts and k sums depend only on (N, boxlen, bins); they come from
// pk_geometry.
//
// A wavefront takes 64 consecutive z' of one row.  Along a row n^2 grows with
// z', so equal bins form contiguous lane runs: a segmented inclusive scan
// (fixed order) leaves each run's total in its last lane, which adds it to
// the wavefront's compensated accumulator of that bin in LDS (different runs,
// different bins, no conflicts).
template <typename T, typename XF>
__global__ void __launch_bounds__(256)
    k_pk_partial(const cx<T>* __restrict__ a, const cx<T>* __restrict__ b, int N, long hs,
                 const int* __restrict__ edges, float cinv, int bins, double* part, XF xf) {
#pragma clang fp contract(off)
    // [4 waves][bins][2] doubles, [bins + 1] bin edges
    extern __shared__ double sm[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    double* acc = sm + (size_t)wv * bins * 2;
    int* first = reinterpret_cast<int*>(sm + (size_t)4 * bins * 2);
    for (int i = lane; i < bins * 2; i += 64) acc[i] = 0;
    for (int i = threadIdx.x; i <= bins; i += blockDim.x) first[i] = edges[i];
    __syncthreads();
    const bool same = (a == b);
    const int H = N / 2 + 1;
    const long nrows = (long)N * N;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    for (long row = (long)blockIdx.x * (blockDim.x >> 6) + wv; row < nrows; row += nwaves) {
        const int x = (int)(row / N), y = (int)(row % N);
        const int nx = freq_index(x, N), ny = freq_index(y, N);
        const long n2xy = (long)nx * nx + (long)ny * ny;
        const double fr = xf.row(x, y);
        const cx<T>* ra = a + row * hs;
        const cx<T>* rb = b + row * hs;
        for (int z0 = 0; z0 < H; z0 += 64) {
            const int z = z0 + lane;
            int bin = -1;
            double v = 0;
            if (z < H) {
                const long n2 = n2xy + (long)z * z;
                if (n2 > 0) {
                    const double f = fr * xf.col(z);
                    const cx<T> u = ra[z];
                    const cx<T> w = same ? u : rb[z];
                    const double mult = (z == 0 || z == N / 2) ? 1.0 : 2.0;
                    bin = pk_bin_of(n2, first, bins, cinv);
                    // (f u).(f w) with the factor applied to each operand
                    v = mult * (((double)u.re * f) * ((double)w.re * f) +
                                ((double)u.im * f) * ((double)w.im * f));
                }
            }
            // segmented inclusive scan over the lane runs of equal bin
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const double vo = __shfl_up(v, o, 64);
                const int bo = __shfl_up(bin, o, 64);
                if (lane >= o && bo == bin) v += vo;
            }
            const int bnext = __shfl_down(bin, 1, 64);
            if (bin >= 0 && (lane == 63 || bnext != bin))
                kadd(acc[bin * 2 + 0], acc[bin * 2 + 1], v);
        }
    }
    __syncthreads();
    for (int bn = threadIdx.x; bn < bins; bn += blockDim.x) {
        double sP = 0, cP = 0;
        for (int w = 0; w < 4; w++) {
            const double* aw = sm + (size_t)w * bins * 2 + bn * 2;
            kadd(sP, cP, aw[0]);
            kadd(sP, cP, aw[1]);
        }
        double* o = part + ((size_t)blockIdx.x * bins + bn) * 2;
        o[0] = sP;
        o[1] = cP;
    }
}

// fold the partials of nblk work-groups in a fixed order: one wave per bin,
// lane l folds work-groups l, l + 64, ... in order, then lane 0 folds the 64
// lane sums in lane order (compensated throughout); out[b*2 + {P_sum,
// K_sum}], obs[b] (the k sums and counts from the geometry tables)
static __global__ void __launch_bounds__(64)
    k_pk_final(const double* part, int nblk, int bins, const double* __restrict__ ksum,
               const long long* __restrict__ kobs, double* out, long long* obs) {
#pragma clang fp contract(off)
    __shared__ double ls[64][2];
    const int bn = blockIdx.x;
    const int lane = threadIdx.x;
    double sP = 0, cP = 0;
    int g = lane;
    for (; g + 3 * 64 < nblk; g += 4 * 64) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double* o = part + ((size_t)(g + u * 64) * bins + bn) * 2;
            v[2 * u] = o[0];
            v[2 * u + 1] = o[1];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) kadd(sP, cP, v[u]);
    }
    for (; g < nblk; g += 64) {
        const double* o = part + ((size_t)g * bins + bn) * 2;
        kadd(sP, cP, o[0]);
        kadd(sP, cP, o[1]);
    }
    ls[lane][0] = sP;
    ls[lane][1] = cP;
    __syncthreads();
    if (lane == 0) {
        double s = 0, c = 0;
        for (int l = 0; l < 64; l++) {
            kadd(s, c, ls[l][0]);
            kadd(s, c, ls[l][1]);
        }
        out[bn * 2 + 0] = s + c;
        out[bn * 2 + 1] = ksum[bn];
        obs[bn] = kobs[bn];
    }
}

// Host: per-bin mode counts (with the multiplicity of src/calc_powerspec.c:
// 75-80) and sums of multiplicity * |k|, as device arrays, cached per
// (device, N, boxlen, bins).  Made from an integer histogram over n^2.
struct PkGeometry {
    const double* ksum;
    const long long* obs;
};
int pk_geometry(int N, double boxlen, int bins, PkGeometry* out);

// Host: device table first[b], b in [0, bins], of the smallest n^2 in
// [1, 3(N/2)^2] whose reference bin is >= b (first[bins] = 3(N/2)^2 + 1).
// Returns nullptr with *err set when the table cannot be made.
const int* pk_bin_edges(int N, double boxlen, int bins, int* err);
// estimate scale of pk_bin_of: (bins-1) / (2 log(sqrt(3) N/2))
inline float pk_bin_cinv(int N, int bins) {
    return (float)((bins - 1) / (2.0 * log(sqrt(3.0) * N / 2)));
}
// Host: number of work-groups used by k_pk_partial
inline int pk_blocks(long nmodes) {
    long b = (nmodes + 255) / 256;
    return (int)(b < 8192 ? b : 8192);
}

// Binned P(k) of two half-complex grids with row stride hs; host outputs
// sums (P, k) and counts.  `xf` transforms each value before the product.
// part_ws/cnt_ws: caller-owned partials (pk_blocks x bins entries), else the
// library workspace.
template <typename T, typename XF>
inline int pk_run(const cx<T>* a, const cx<T>* b, int N, long hs, double boxlen, int bins, XF xf,
                  double* sums, long long* obs, hipStream_t s, double* dev_sums,
                  long long* dev_obs, double* part_ws = nullptr, long long* cnt_ws = nullptr) {
    int err = MITOS_OK;
    const int* edges = pk_bin_edges(N, boxlen, bins, &err);
    if (!edges) return err;
    PkGeometry geo;
    MITOS_TRY(pk_geometry(N, boxlen, bins, &geo));
    const long nmodes = (long)N * N * (N / 2 + 1);
    const int nblk = pk_blocks(nmodes);
    const size_t pbytes = sizeof(double) * 2 * bins * nblk;
    char* ws = nullptr;
    if (!part_ws || !dev_sums) {
        ws = (char*)workspace(2, pbytes + sizeof(double) * 2 * bins + sizeof(long long) * bins);
        if (!ws) return MITOS_ERR_ALLOC;
    }
    (void)cnt_ws;
    double* part = part_ws ? part_ws : (double*)ws;
    double* fs = dev_sums ? dev_sums : (double*)(ws + pbytes);
    long long* fo = dev_obs ? dev_obs : (long long*)(ws + pbytes + sizeof(double) * 2 * bins);
    const size_t shm = sizeof(double) * 4 * 2 * bins + sizeof(int) * (bins + 1);
    hipLaunchKernelGGL((k_pk_partial<T, XF>), dim3(nblk), dim3(256), shm, s, a, b, N, hs, edges,
                       pk_bin_cinv(N, bins), bins, part, xf);
    MITOS_TRY(launch_status());
    hipLaunchKernelGGL(k_pk_final, dim3(bins), dim3(64), 0, s, part, nblk, bins,
                       geo.ksum, geo.obs, fs, fo);
    MITOS_TRY(launch_status());
    if (sums) {
        MITOS_HIP_CHECK(hipMemcpyAsync(sums, fs, sizeof(double) * 2 * bins,
                                       hipMemcpyDeviceToHost, s));
        MITOS_HIP_CHECK(hipMemcpyAsync(obs, fo, sizeof(long long) * bins, hipMemcpyDeviceToHost,
                                       s));
        MITOS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return MITOS_OK;
}

// src/calc_powerspec.c:84-89 on the host: averages, empty bins give 0/0 = NaN
void pk_finish(const double* sums, const long long* obs, int bins, double boxlen, double* k_out,
               double* p_out);

}  // namespace mitos
