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
// A wavefront takes one row (x, y) at a time; lane l takes the C = ceil(H/64)
// consecutive z' in [l C, (l + 1) C).  Along a row n^2 grows with z', so the
// modes of one bin form one run.  Each lane sums its values run by run in
// order: a run that starts and ends inside the lane is the whole bin of this
// row and goes straight to the wavefront's compensated (Neumaier) accumulator
// of that bin in LDS; the lane's first run (head) and last run (tail) may
// continue in the neighbouring lanes.  A tail is added to the next lane's
// head when their bins agree, then a segmented inclusive scan (fixed order)
// over the heads leaves each run's total in its last lane, which adds it to
// the accumulator.  Every bin of a row is added once, so lanes never write
// the same accumulator at once, and the order of every sum is fixed.
// The lane's z', z'^2, multiplicities and column factors are fixed across
// rows and held in registers; the next row's values are loaded while the
// current one is summed.  CM >= C (compile time); CM = 0: any C, no
// prefetch, and a cross spectrum of two grids.
template <typename T, typename XF, int CM>
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
    const int C = (H + 63) / 64;
    const int z0 = lane * C, z1 = z0 + C < H ? z0 + C : H;
    const int nrows = N * N;
    const int nwaves = (int)gridDim.x * (int)(blockDim.x >> 6);
    // rows of this wavefront: row0, row0 + nwaves, ...; (x, y) advance by
    // (nwaves / N, nwaves % N) with a carry
    const int row0 = (int)blockIdx.x * (int)(blockDim.x >> 6) + wv;
    const int dxr = nwaves / N, dyr = nwaves - dxr * N;
    int x = row0 / N, y = row0 - x * N;
    // one row: v(z) summed run by run, then merged across the lanes
    auto row_sum = [&](int xx, int yy, auto value_at) {
        const int nx = freq_index(xx, N), ny = freq_index(yy, N);
        const int n2xy = nx * nx + ny * ny;
        // this lane's runs: head (hb, hsum) once a second run starts, the
        // current run (cb, cs)
        int hb = -1, cb = -1;
        double hsum = 0, cs = 0;
        auto push = [&](int n2, double v) {
            int bn;
            if (cb < 0) {
                bn = pk_bin_of(n2, first, bins, cinv);
            } else {
                bn = cb;
                while (bn < bins - 1 && n2 >= first[bn + 1]) bn++;
            }
            if (bn == cb) {
                cs += v;
            } else {
                if (cb >= 0) {
                    if (hb < 0) {
                        hb = cb;
                        hsum = cs;
                    } else {
                        kadd(acc[cb * 2 + 0], acc[cb * 2 + 1], cs);  // a run inside the lane
                    }
                }
                cb = bn;
                cs = v;
            }
        };
        value_at(n2xy, push);
        const bool multi = hb >= 0;
        const int headb = multi ? hb : cb;
        double head = multi ? hsum : cs;
        const int tailb = multi ? cb : -1;
        // the previous lane's tail joins this head when the bins agree; a
        // tail that does not continue is a whole run
        const int pb = __shfl_up(tailb, 1, 64);
        const double ps = __shfl_up(multi ? cs : 0.0, 1, 64);
        const int nb = __shfl_down(headb, 1, 64);
        if (lane > 0 && pb >= 0 && pb == headb) head += ps;
        if (multi && !(lane < 63 && nb == tailb)) kadd(acc[tailb * 2 + 0], acc[tailb * 2 + 1], cs);
        // segmented inclusive scan over the heads of equal bin
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double vo = __shfl_up(head, o, 64);
            const int bo = __shfl_up(headb, o, 64);
            if (lane >= o && bo == headb) head += vo;
        }
        if (headb >= 0 && (lane == 63 || nb != headb))
            kadd(acc[headb * 2 + 0], acc[headb * 2 + 1], head);
    };
    if constexpr (CM > 0) {
        // per-lane constants of the lane's z'
        int zsq[CM];
        double colm[CM];  // multiplicity (src/calc_powerspec.c:75-80) / column factor
        double colf[CM];
#pragma unroll
        for (int k = 0; k < CM; k++) {
            const int z = z0 + k;
            zsq[k] = z * z;
            colf[k] = (k < C && z < H) ? xf.col(z) : 0.0;
            colm[k] = (z == 0 || z == N / 2) ? 1.0 : 2.0;
        }
        const int nk = z1 > z0 ? z1 - z0 : 0;
        cx<T> cur[CM];
        const cx<T>* rp = a + (long)row0 * hs + z0;
#pragma unroll
        for (int k = 0; k < CM; k++)
            if (k < nk && row0 < nrows) cur[k] = rp[k];
        for (int row = row0; row < nrows; row += nwaves) {
            // the next row's values and factor first
            const int rn = row + nwaves;
            int xn = x + dxr, yn = y + dyr;
            if (yn >= N) {
                yn -= N;
                xn++;
            }
            cx<T> nxt[CM];
            const cx<T>* np = a + (long)rn * hs + z0;
#pragma unroll
            for (int k = 0; k < CM; k++)
                if (k < nk && rn < nrows) nxt[k] = np[k];
            const double fr = xf.row(x, y);
            row_sum(x, y, [&](int n2xy, auto& push) {
#pragma unroll
                for (int k = 0; k < CM; k++) {
                    if (k >= nk) break;
                    const int n2 = n2xy + zsq[k];
                    if (n2 == 0) continue;  // the DC mode
                    const double f = fr * colf[k];
                    const double ur = (double)cur[k].re * f, ui = (double)cur[k].im * f;
                    // (f u).(f u) with the factor applied to each operand
                    push(n2, colm[k] * (ur * ur + ui * ui));
                }
            });
#pragma unroll
            for (int k = 0; k < CM; k++) cur[k] = nxt[k];
            x = xn;
            y = yn;
        }
    } else {
        for (int row = row0; row < nrows; row += nwaves) {
            const double fr = xf.row(x, y);
            const cx<T>* ra = a + (long)row * hs;
            const cx<T>* rb = b + (long)row * hs;
            row_sum(x, y, [&](int n2xy, auto& push) {
                for (int z = z0; z < z1; z++) {
                    const int n2 = n2xy + z * z;
                    if (n2 == 0) continue;  // the DC mode
                    const double f = fr * xf.col(z);
                    const cx<T> u = ra[z];
                    const cx<T> w = same ? u : rb[z];
                    const double mult = (z == 0 || z == N / 2) ? 1.0 : 2.0;
                    // (f u).(f w) with the factor applied to each operand
                    push(n2, mult * (((double)u.re * f) * ((double)w.re * f) +
                                     ((double)u.im * f) * ((double)w.im * f)));
                }
            });
            x += dxr;
            y += dyr;
            if (y >= N) {
                y -= N;
                x++;
            }
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
    const int C = (N / 2 + 64) / 64;
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nblk), dim3(256), shm, s, a, b, N, hs, edges,
                           pk_bin_cinv(N, bins), bins, part, xf);
    };
    if (a != b || C > 16)
        launch(k_pk_partial<T, XF, 0>);
    else if (C <= 4)
        launch(k_pk_partial<T, XF, 4>);
    else if (C <= 8)
        launch(k_pk_partial<T, XF, 8>);
    else if (C <= 12)
        launch(k_pk_partial<T, XF, 12>);
    else
        launch(k_pk_partial<T, XF, 16>);
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
