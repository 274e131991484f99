// Gaussian random field: white noise from the xoshiro256** stream in the
// reference's draw order, with a k-dependent amplitude fused in.
#pragma once

#include "crmath.h"
#include "rng.h"

namespace mitos {

// Per-grid-size jump data, resident on the device.
struct NoiseJumps {
    const uint64_t* tab;  // nbits matrices J^(2^b), J = T^(4*(N/2+1)), 256*4 words each
    int nbits;
    Xo s0;    // state before the first draw (row 0 starts here)
    Xo base;  // state after 4*(N/2+1) - 4 draws (row 1 starts here)
};

// Host: jump data for grid size N starting from state s0 (table cached per N).
int noise_jumps(int N, const Xo& s0, NoiseJumps* out);
// Host: state after the whole grid's draws, 4*(N*N*(N/2+1) - 1).
int noise_advance(int N, Xo* st);

// x * 2^s + x, written so that it stays one shift-add (v_lshl_add_u64): the
// empty asm hides that both operands are x, so the compiler cannot fold the
// sum back into a 64-bit multiply (two v_mad_u64_u32)
__device__ __forceinline__ uint64_t shl_add64(uint64_t x, int s) {
    uint64_t y = x;
    asm("" : "+v"(y));
    return (y << s) + x;
}

// 64-bit rotations and the shift by 17 on 32-bit halves with v_alignbit_b32
// (alignbit(a, b, c) = low word of ((a:b) >> c)): two instructions each
template <int K>
__device__ __forceinline__ uint64_t rotl64_dev(uint64_t x) {
    static_assert(K > 0 && K < 64 && K != 32, "rotation");
    const unsigned lo = (unsigned)x, hi = (unsigned)(x >> 32);
    unsigned nlo, nhi;
    if constexpr (K < 32) {
        nhi = __builtin_amdgcn_alignbit(hi, lo, 32 - K);
        nlo = __builtin_amdgcn_alignbit(lo, hi, 32 - K);
    } else {
        nhi = __builtin_amdgcn_alignbit(lo, hi, 64 - K);
        nlo = __builtin_amdgcn_alignbit(hi, lo, 64 - K);
    }
    return ((uint64_t)nhi << 32) | nlo;
}
template <int K>
__device__ __forceinline__ uint64_t shl64_dev(uint64_t x) {
    static_assert(K > 0 && K < 32, "shift");
    const unsigned lo = (unsigned)x, hi = (unsigned)(x >> 32);
    return ((uint64_t)__builtin_amdgcn_alignbit(hi, lo, 32 - K) << 32) | (lo << K);
}

// xo_next on the device: the same draw, with s1 * 5 and r * 9 as shift-adds
// and the rotations and shift on 32-bit halves
__device__ __forceinline__ uint64_t xo_next_dev(Xo& st) {
    const uint64_t out = shl_add64(rotl64_dev<7>(shl_add64(st.s[1], 2)), 3);
    const uint64_t t = shl64_dev<17>(st.s[1]);
    st.s[2] ^= st.s[0];
    st.s[3] ^= st.s[1];
    st.s[1] ^= st.s[2];
    st.s[0] ^= st.s[3];
    st.s[2] ^= t;
    st.s[3] = rotl64_dev<45>(st.s[3]);
    return out;
}

// Start states of the 256 z-rows line0 + t of a 256-thread work-group (line0
// a multiple of 256), by the whole work-group.  A product M v of a jump
// matrix and a state takes one parity per thread (output bit t = parity of
// row t of M and v) and a ballot per wave (wave w holds word w).  The
// work-group advances J.base to row line0 (the matrices of the bits of
// line0 - 256 above bit 7, then J^255), then doubles: st[j + 2^k] =
// J^(2^k) st[j] for k < 8, so thread t reads its row's state st[t].  That is
// 255 + up to 20 products shared by 256 rows, against up to 20 per row when
// every thread jumps on its own.  lds: 256 * 4 + 4 words.
__device__ __forceinline__ Xo block_row_start(const NoiseJumps& J, long line0, uint64_t* lds) {
    uint64_t(*st)[4] = reinterpret_cast<uint64_t(*)[4]>(lds);
    uint64_t* xw = lds + 256 * 4;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    auto apply = [&](int b, const Xo& v) -> Xo {
        const uint64_t* row = J.tab + (size_t)b * 1024 + (size_t)t * 4;
        const uint64_t p = (row[0] & v.s[0]) ^ (row[1] & v.s[1]) ^ (row[2] & v.s[2]) ^
                           (row[3] & v.s[3]);
        const uint64_t w = __ballot(__builtin_popcountll(p) & 1);
        if (lane == 0) xw[wv] = w;
        __syncthreads();
        const Xo r = {{xw[0], xw[1], xw[2], xw[3]}};
        __syncthreads();
        return r;
    };
    Xo v = J.base;
    if (line0 > 0) {
        // row line0 starts after line0 - 1 jumps: (line0 - 256) + 255
        const unsigned long hi = (unsigned long)(line0 - 256);
        for (int b = 8; b < J.nbits; b++)
            if ((hi >> b) & 1UL) v = apply(b, v);
        for (int b = 0; b < 8; b++) v = apply(b, v);
    }
    if (t < 4) st[0][t] = v.s[t];
    __syncthreads();
    for (int k = 0; k < 8 && k < J.nbits; k++) {
        const int m = 1 << k;
        const uint64_t* row = J.tab + (size_t)k * 1024 + (size_t)t * 4;
        const uint64_t r0 = row[0], r1 = row[1], r2 = row[2], r3 = row[3];
        for (int j = 0; j < m; j++) {
            const uint64_t p = (r0 & st[j][0]) ^ (r1 & st[j][1]) ^ (r2 & st[j][2]) ^
                               (r3 & st[j][3]);
            const uint64_t w = __ballot(__builtin_popcountll(p) & 1);
            if (lane == 0) st[j + m][wv] = w;
        }
        __syncthreads();
    }
    // row 0 starts at s0 (the DC mode takes no draws); row line starts after
    // line - 1 jumps from J.base
    const int j = line0 > 0 ? t : t - 1;
    Xo r = J.s0;
    if (j >= 0) r = Xo{{st[j][0], st[j][1], st[j][2], st[j][3]}};
    __syncthreads();  // the table's reads precede the caller's reuse of lds
    return r;
}

// src/random.c:34-49: cosine branch of Box-Muller from two 64-bit draws.
// CR: correctly rounded log/cos (bit-exact Gaussians); else the device libm
// (within an ulp, used when the result is rounded to float anyway).
template <bool CR = true>
__device__ __forceinline__ double bm_normal(Xo& gen) {
    const uint64_t A = xo_next_dev(gen);
    const uint64_t B = xo_next_dev(gen);
    const double RMax = 18446744073709551616.0;  // (double)UINT64_MAX + 1 = 2^64
    const double u = ((double)A + 0.5) / RMax;
    const double v = ((double)B + 0.5) / RMax;
    if (CR) return sqrt(-2 * cr_log(u)) * cr_cos(2 * M_PI * v);
    return sqrt(-2 * log(u)) * cos(2 * M_PI * v);
}

// cos(2 pi v) in float for v in [0, 1): r = v - rint(v) is exact, t = |r|
// (or 1/2 - |r| past a quarter period, with the sign flipped) lies in
// [0, 1/4], and cos(2 pi t) is a degree-10 even polynomial in t fitted here
// (weighted least squares on Chebyshev nodes, iterated towards equal ripple;
// fit error 2.2e-10, float evaluation within 1.4e-7 of cos(2 pi v))
__device__ __forceinline__ float cos2pi_f(float v) {
    const float r = v - rintf(v);
    const float s = fabsf(r);
    const bool far = s > 0.25f;
    const float t = far ? 0.5f - s : s;
    const float t2 = t * t;
    float p = -24.982326173851263f;
    p = fmaf(p, t2, 60.14403025957297f);
    p = fmaf(p, t2, -85.45357257295981f);
    p = fmaf(p, t2, 64.9393466552905f);
    p = fmaf(p, t2, -19.73920854920032f);
    p = fmaf(p, t2, 0.9999999997815415f);
    return far ? -p : p;
}

// The Box-Muller pair of bm_normal in float arithmetic (FP32 pipeline): the
// same draws; the 64-bit words become floats as hi * 2^32 + lo with one
// rounding (within an ulp of the rounded word), then log, the hardware square
// root (an ulp) and cos2pi_f.
__device__ __forceinline__ float u64_to_float(uint64_t A) {
    // the empty asm keeps the high word a 32-bit value, so its conversion is
    // one v_cvt_f32_u32 rather than a 64-bit integer conversion
    unsigned hi = (unsigned)(A >> 32);
    asm("" : "+v"(hi));
    return fmaf((float)hi, 4294967296.0f, (float)(unsigned)A);
}
__device__ __forceinline__ float bm_normal_f(Xo& gen) {
    const uint64_t A = xo_next_dev(gen);
    const uint64_t B = xo_next_dev(gen);
    const float s = 5.42101086242752217e-20f;  // 2^-64
    const float u = (u64_to_float(A) + 0.5f) * s;
    const float v = (u64_to_float(B) + 0.5f) * s;
    return __builtin_amdgcn_sqrtf(-2.0f * logf(u)) * cos2pi_f(v);
}

// src/fft.c:29-35, evaluated without contraction so k matches the host bits
__host__ __device__ __forceinline__ void wavevec(int x, int y, int z, int N, double dk,
                                                 double& kx, double& ky, double& kz,
                                                 double& k) {
#pragma clang fp contract(off)
    kx = (x > N / 2) ? (x - N) * dk : x * dk;
    ky = (y > N / 2) ? (y - N) * dk : y * dk;
    kz = (z > N / 2) ? (z - N) * dk : z * dk;
    k = sqrt(kx * kx + ky * ky + kz * kz);
}

// src/grf.c:26-69 with an amplitude amp(n2) applied to both parts (n2 =
// integer |k|^2 / dk^2).  One thread walks one z-row in stream
// order; a block's 256 rows are staged through LDS in chunks of 16 modes so
// the stores are 16-mode contiguous runs.
template <typename T, typename AMP, bool CR>
__global__ void __launch_bounds__(256)
    k_grf(cx<T>* out, long hs, int N, double boxlen, long batch_stride, NoiseJumps J,
          AMP amp) {
    constexpr int ZC = 16;
    __shared__ cx<T> stage[256 * ZC];
    const long nrows = (long)N * N;
    const long line0 = (long)blockIdx.x * 256;
    const long line = line0 + threadIdx.x;
    const bool ok = line < nrows;
    const int x = (int)(line / N), y = (int)(line % N);
    Xo gen = block_row_start(J, line0, reinterpret_cast<uint64_t*>(stage));
    const double boxvol = boxlen * boxlen * boxlen;
    const double factor = sqrt(boxvol / 2);
    const int H = N / 2 + 1;
    // stored z': the modes, and the row padding up to a whole chunk (zeros),
    // so the last chunk writes whole 128-byte lines
    const long hz = (long)(H + ZC - 1) / ZC * ZC;
    const int WZ = (int)(hz < hs ? hz : hs);
    const int nx = freq_index(x, N), ny = freq_index(y, N);
    const long n2xy = (long)nx * nx + (long)ny * ny;
    cx<T>* base = out + blockIdx.y * batch_stride;
    for (int z0 = 0; z0 < H; z0 += ZC) {
        if (ok) {
            for (int dz = 0; dz < ZC; dz++) {
                const int z = z0 + dz;
                // past the last mode: zeros for the row padding
                if (z >= H) {
                    stage[threadIdx.x * ZC + dz] = mk<T>(0, 0);
                    continue;
                }
                // the DC mode (n^2 = 0, k = 0) takes no draws
                const long n2 = n2xy + (long)z * z;
                cx<T> val = mk<T>(0, 0);
                if (n2 > 0) {
                    if constexpr (sizeof(T) == 4) {
                        const float ga = bm_normal_f(gen);
                        const float gb = bm_normal_f(gen);
                        float fm;
                        if constexpr (AMP::prescaled)
                            fm = amp.f(n2);  // (float)(factor * amp) made on the host
                        else
                            fm = (float)(factor * amp(n2));
                        val = mk<T>(ga * fm, gb * fm);
                    } else {
                        const double a = bm_normal<CR>(gen) * factor;
                        const double b = bm_normal<CR>(gen) * factor;
                        const double m = amp(n2);
                        val = mk<T>((T)(a * m), (T)(b * m));
                    }
                }
                stage[threadIdx.x * ZC + dz] = val;
            }
        }
        __syncthreads();
        for (int i = 0; i < ZC; i++) {
            const int idx = i * 256 + threadIdx.x;
            const int r = idx / ZC, dz = idx % ZC;
            const long l = line0 + r;
            if (l < nrows && z0 + dz < WZ) st(base + l * hs + z0 + dz, stage[r * ZC + dz]);
        }
        __syncthreads();
    }
}

// src/grf.c:78-175 on one rank.  In place: every lower-half point reads its
// mirror, which lies in the untouched upper half or is the point itself.
template <typename T>
__global__ void k_hermitian(cx<T>* f, long hs, int N, long batch_stride) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)(N / 2 + 1) * N;
    if (i >= 2 * per) return;
    const int z = (i < per) ? 0 : N / 2;
    const long j = i % per;
    const int x = (int)(j / N), y = (int)(j % N);
    if ((x == 0 || x == N / 2) && y > N / 2) return;
    const int invx = (x > 0) ? N - x : 0;
    const int invy = (y > 0) ? N - y : 0;
    const int invz = (z > 0) ? N - z : 0;
    cx<T>* g = f + blockIdx.y * batch_stride;
    const long id = ((long)x * N + y) * hs + z;
    if (invx == x && invy == y && invz == z) {
        g[id].im = 0;
    } else {
        const long mid = ((long)invx * N + invy) * hs + z;
        g[id] = conjg(g[mid]);
    }
}

struct AmpOne {
    static constexpr bool prescaled = false;
    __device__ __forceinline__ double operator()(long) const { return 1.0; }
};

// amplitude looked up by the integer n^2 of the mode
struct AmpTable {
    static constexpr bool prescaled = false;
    const double* tab;
    __device__ __forceinline__ double operator()(long n2) const { return tab[n2]; }
};

// FP32 noise: the amplitude times the noise factor sqrt(L^3/2), rounded to
// float on the host, so the values are those of AmpTable in half the bytes
// (the table of a 1024^3 grid, 3.1 MB, fits one XCD's L2)
struct AmpTableF {
    static constexpr bool prescaled = true;
    const float* tabf;
    __device__ __forceinline__ float f(long n2) const { return tabf[n2]; }
    __device__ __forceinline__ double operator()(long n2) const { return tabf[n2]; }
};

template <typename T, typename AMP, bool CR = true>
int launch_grf(cx<T>* out, long hs, int N, double boxlen, int nb, long bstride,
               const NoiseJumps& J, AMP amp, hipStream_t s) {
    const long rows = (long)N * N;
    dim3 grid((unsigned)((rows + 255) / 256), nb);
    hipLaunchKernelGGL((k_grf<T, AMP, CR>), grid, dim3(256), 0, s, out, hs, N, boxlen, bstride,
                       J, amp);
    return hipGetLastError() == hipSuccess ? MITOS_OK : MITOS_ERR_DEVICE;
}

template <typename T>
int launch_hermitian(cx<T>* f, long hs, int N, int nb, long bstride, hipStream_t s) {
    const long n = 2L * (N / 2 + 1) * N;
    dim3 grid((unsigned)((n + 255) / 256), nb);
    hipLaunchKernelGGL((k_hermitian<T>), grid, dim3(256), 0, s, f, hs, N, bstride);
    return hipGetLastError() == hipSuccess ? MITOS_OK : MITOS_ERR_DEVICE;
}

}  // namespace mitos
