// 3D real<->complex FFT passes over the half-complex grid layout.
//
//  strided pass : c2c lines along x (stride N*hs) or y (stride hs).  A
//                 work-group takes C consecutive z' columns at one value of the
//                 other coordinate, so every load/store instruction of a wave
//                 covers rows of C contiguous complex values.
//  z c2r pass   : one row of N/2+1 complex values -> N reals, through an
//                 N/2-point complex FFT (even/odd packing).  The imaginary
//                 parts of the z'=0 and z'=N/2 entries are ignored, which is
//                 how a c2r treats a non-Hermitian last-axis spectrum.
//  z r2c pass   : N reals -> N/2+1 complex values, same packing.
//
// Each pass takes an operator object that may rewrite a value on load (pre)
// or before the store (post) given its grid coordinates, so k-space kernels
// and normalisations fuse into the passes.
#pragma once

#include "fft_core.h"

namespace mitos {

template <int... Rs>
struct Radices {};

// line configuration per length: E values per thread and the stage radices
template <int L>
struct LineCfg;
// (lengths 3*2^k hold 12 values per thread: radices 4, 2 and a final 3, so
// every line of the 3*2^k grids has a power-of-two thread count)
template <> struct LineCfg<4>    { static constexpr int E = 4;  typedef Radices<4> R; };
template <> struct LineCfg<8>    { static constexpr int E = 8;  typedef Radices<8> R; };
template <> struct LineCfg<12>   { static constexpr int E = 12; typedef Radices<4, 3> R; };
template <> struct LineCfg<16>   { static constexpr int E = 16; typedef Radices<16> R; };
template <> struct LineCfg<24>   { static constexpr int E = 12; typedef Radices<4, 2, 3> R; };
template <> struct LineCfg<32>   { static constexpr int E = 8;  typedef Radices<8, 4> R; };
template <> struct LineCfg<48>   { static constexpr int E = 12; typedef Radices<4, 4, 3> R; };
template <> struct LineCfg<64>   { static constexpr int E = 8;  typedef Radices<8, 8> R; };
template <> struct LineCfg<96>   { static constexpr int E = 12; typedef Radices<4, 4, 2, 3> R; };
template <> struct LineCfg<128>  { static constexpr int E = 16; typedef Radices<16, 8> R; };
template <> struct LineCfg<192>  { static constexpr int E = 12; typedef Radices<4, 4, 4, 3> R; };
template <> struct LineCfg<256>  { static constexpr int E = 16; typedef Radices<16, 16> R; };
template <> struct LineCfg<384>  { static constexpr int E = 12; typedef Radices<4, 4, 4, 2, 3> R; };
template <> struct LineCfg<512>  { static constexpr int E = 8;  typedef Radices<8, 8, 8> R; };
template <> struct LineCfg<768>  { static constexpr int E = 12; typedef Radices<4, 4, 4, 4, 3> R; };
template <> struct LineCfg<1024> { static constexpr int E = 16; typedef Radices<16, 16, 4> R; };
template <> struct LineCfg<1536> { static constexpr int E = 24; typedef Radices<8, 8, 8, 3> R; };

// Line configuration with fewer values per thread, for kernels that keep
// several transforms' worth of data live (multi-output passes, fused z pass).
template <int L>
struct LineCfgS;
template <> struct LineCfgS<4>    { static constexpr int E = 4;  typedef Radices<4> R; };
template <> struct LineCfgS<8>    { static constexpr int E = 8;  typedef Radices<8> R; };
template <> struct LineCfgS<12>   { static constexpr int E = 12; typedef Radices<4, 3> R; };
template <> struct LineCfgS<16>   { static constexpr int E = 4;  typedef Radices<4, 4> R; };
template <> struct LineCfgS<24>   { static constexpr int E = 12; typedef Radices<4, 2, 3> R; };
template <> struct LineCfgS<32>   { static constexpr int E = 8;  typedef Radices<8, 4> R; };
template <> struct LineCfgS<48>   { static constexpr int E = 12; typedef Radices<4, 4, 3> R; };
template <> struct LineCfgS<64>   { static constexpr int E = 8;  typedef Radices<8, 8> R; };
template <> struct LineCfgS<96>   { static constexpr int E = 12; typedef Radices<4, 4, 2, 3> R; };
template <> struct LineCfgS<128>  { static constexpr int E = 8;  typedef Radices<8, 8, 2> R; };
template <> struct LineCfgS<192>  { static constexpr int E = 12; typedef Radices<4, 4, 4, 3> R; };
template <> struct LineCfgS<256>  { static constexpr int E = 8;  typedef Radices<8, 8, 4> R; };
template <> struct LineCfgS<384>  { static constexpr int E = 12; typedef Radices<4, 4, 4, 2, 3> R; };
template <> struct LineCfgS<512>  { static constexpr int E = 8;  typedef Radices<8, 8, 8> R; };
template <> struct LineCfgS<768>  { static constexpr int E = 12; typedef Radices<4, 4, 4, 4, 3> R; };
template <> struct LineCfgS<1024> { static constexpr int E = 8;  typedef Radices<8, 8, 8, 2> R; };
template <> struct LineCfgS<1536> { static constexpr int E = 12; typedef Radices<4, 4, 4, 4, 2, 3> R; };

// columns per strided tile: 128-byte rows (16 float2 / 8 double2)
template <typename T>
struct ColTile {
    static constexpr int C = 128 / (2 * sizeof(T));
};

// ---- operators ------------------------------------------------------------
// pre/post(v, a, b, c): (a, b, c) are the (x, y, z') grid coordinates of the
// value in half-complex index space.
struct OpNone {
    template <typename T>
    __device__ __forceinline__ cx<T> pre(cx<T> v, int, int, int) const { return v; }
    template <typename T>
    __device__ __forceinline__ cx<T> post(cx<T> v, int, int, int) const { return v; }
};

template <typename T>
struct OpScale {
    T s;
    __device__ __forceinline__ cx<T> pre(cx<T> v, int, int, int) const { return v; }
    __device__ __forceinline__ cx<T> post(cx<T> v, int, int, int) const { return cscale(v, s); }
};

// Columns a strided pass stores: the valid ones rounded up to whole 128-byte
// lines (or tiles of C, if wider) while the row padding allows.  The last
// tiles then write whole lines (zeros in the padding columns) instead of one
// partial line per row, which the memory system completes with an extra read.
// A launch of stored_cols / C tiles writes every line whole.
template <typename T, int C>
__host__ __device__ __forceinline__ int stored_cols(int ncols, long hs) {
    constexpr int W = C * (int)sizeof(cx<T>) >= 128 ? C : 128 / (int)sizeof(cx<T>);
    const long r = (long)(ncols + W - 1) / W * W;
    return (int)(r < hs ? r : hs);
}

// ---- strided c2c pass ------------------------------------------------------
// AXIS 0: lines along x, outer index o = y;  AXIS 1: lines along y, o = x.
// Work-group (bx, by) transforms column tile bx for the KO outer indices
// o = o0 + by*KO + i, i < KO (consecutive y rows of an x pass stream faster
// from one work-group than from KO); blockIdx.z selects one of several grids
// spaced `batch_stride` apart.  in == out is allowed: every thread stores
// exactly the elements it loaded.
template <typename T, int L, int E, int C, int AXIS, int DIR, bool SPLIT, int KO, typename OP,
          int... Rs>
__global__ void __launch_bounds__(L / E * C)
    k_pass_strided(const cx<T>* in, cx<T>* out, long hs, int ncols, long batch_stride,
                   const cx<T>* __restrict__ tw, OP op, int o0) {
    constexpr int TPL = L / E;
    typedef LdsMap<T, L, C, true> M;
    __shared__ T lds[(SPLIT ? 1 : 2) * M::SIZE];
    const int c = threadIdx.x % C;
    const int t = threadIdx.x / C;
    unsigned bx, by;
    xcd_block2(bx, by);
    const int col = bx * C + c;
    const bool ok = col < ncols;
    const bool wok = col < stored_cols<T, C>(ncols, hs);
    const long pstride = (AXIS == 0) ? (long)L * hs : hs;
    const long ostride = (AXIS == 0) ? hs : (long)L * hs;
    // loads address a work-group-uniform base per element plus one 32-bit
    // per-thread offset (the launcher checks TPL * pstride elements fit in 32
    // bits): no 64-bit address registers per element
    const unsigned toff = (unsigned)(((long)t * pstride + c) * (long)sizeof(cx<T>));
#pragma unroll 1
    for (int i = 0; i < KO; i++) {
        const int o = o0 + (int)by * KO + i;
        const long base = blockIdx.z * batch_stride + (long)o * ostride + col;
        const char* ub = reinterpret_cast<const char*>(in + blockIdx.z * batch_stride +
                                                       (long)o * ostride + (long)bx * C);
        // per-row copies: LDS addresses and twiddles kept live across the rows
        // would take registers from the values
        int tl = t, cl = c;
        const cx<T>* twl = tw;
        if (KO > 1) asm volatile("" : "+v"(tl), "+v"(cl), "+s"(twl));
        cx<T> v[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int p = tl + TPL * e;
            const char* eb = ub + (long)TPL * e * pstride * (long)sizeof(cx<T>);
            // columns past ncols (tile and padding) are zeros; the operator
            // (which may index tables by column) sees valid columns only
            cx<T> a = ok ? ld(reinterpret_cast<const cx<T>*>(eb + toff)) : mk<T>(0, 0);
            if (ok) a = (AXIS == 0) ? op.pre(a, p, o, col) : op.pre(a, o, p, col);
            v[e] = a;
        }
        line_fft<T, L, E, C, true, DIR, Rs...>(v, lds, SPLIT ? lds : lds + M::SIZE, twl, tl, cl);
        if (wok) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int p = tl + TPL * e;
                cx<T> a = !ok ? mk<T>(0, 0)
                              : (AXIS == 0) ? op.post(v[e], p, o, col) : op.post(v[e], o, p, col);
                st(out + base + (long)p * pstride, a);
            }
        }
    }
}

// ---- z c2r pass ------------------------------------------------------------
// in: rows of hs complex (first N/2+1 used); out: rows of rs reals.  HO = 0:
// the N values at row[0..N), the remaining rs-N set to zero.  HO > 0: value z
// at row[HO + z] with the periodic halo row[HO-1] = value N-1 and row[HO+N] =
// value 0 (rs >= N + HO + 1; the rest of the row is set to zero).  L = N/2.
// Rows r0 .. r1-1 are transformed.  Each work-group walks row groups
// blockIdx.x, blockIdx.x + gridDim.x, ... and loads the next group's row while
// transforming the current one.
template <typename T, int L, int E, int RW, int HO, typename OP, int... Rs>
__global__ void __launch_bounds__(L / E * RW)
    k_zc2r(const cx<T>* __restrict__ in, T* __restrict__ out, long r0, long r1, int N, long hs,
           long rs, long in_batch, long out_batch, const cx<T>* __restrict__ tw0,
           const cx<T>* __restrict__ twN0, T scale, OP op) {
    constexpr int TPL = L / E;
    typedef LdsMap<T, L, RW, false> M;
    __shared__ T lre[M::SIZE];
    __shared__ T lim[M::SIZE];
    const int t0 = threadIdx.x % TPL;
    const int rl0 = threadIdx.x / TPL;
    const long ngroups = (r1 - r0 + RW - 1) / RW;
    cx<T> nxt[E + 1];
    auto load = [&](long g) {
        const long row = r0 + g * RW + rl0;
        if (row < r1) {
            const cx<T>* src = in + blockIdx.y * in_batch + row * hs;
#pragma unroll
            for (int e = 0; e < E; e++) nxt[e] = ld(src + t0 + TPL * e);
            nxt[E] = t0 == 0 ? ld(src + L) : mk<T>(0, 0);
        }
    };
    if ((long)blockIdx.x < ngroups) load(blockIdx.x);
    for (long g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int t = t0, rl = rl0;
        const cx<T>* tw = tw0;
        const cx<T>* twN = twN0;
        const long row = r0 + g * RW + rl;
        const bool ok = row < r1;
        const int gx = (int)(row / N), gy = (int)(row % N);
        // stage the row (positions 0..L) in LDS
        if (ok) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = t + TPL * e;
                cx<T> a = op.pre(nxt[e], gx, gy, k);
                lre[M::idx(k, rl)] = a.re;
                lim[M::idx(k, rl)] = a.im;
            }
            if (t == 0) {
                cx<T> a = op.pre(nxt[E], gx, gy, L);
                lre[M::idx(L, rl)] = a.re;
                lim[M::idx(L, rl)] = a.im;
            }
        }
        if (g + gridDim.x < ngroups) load(g + gridDim.x);
        __syncthreads();
        cx<T> v[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = t + TPL * e;
            cx<T> xk = mk<T>(lre[M::idx(k, rl)], lim[M::idx(k, rl)]);
            cx<T> xm = mk<T>(lre[M::idx(L - k, rl)], lim[M::idx(L - k, rl)]);
            if (k == 0) {  // Im of the DC and Nyquist entries is ignored
                xk.im = 0;
                xm.im = 0;
            }
            cx<T> a = xk + conjg(xm);
            cx<T> b = xk - conjg(xm);
            cx<T> w = twN[k];
            w.im = -w.im;  // exp(+2 pi i k / N)
            v[e] = a + mul_i<1>(cmul(w, b));
        }
        __syncthreads();
        line_fft<T, L, E, RW, false, 1, Rs...>(v, lre, lim, tw, t, rl);
        if (ok) {
            T* dst = out + blockIdx.y * out_batch + row * rs;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int m = t + TPL * e;
                typename vec2<T>::type p;
                p.x = v[e].re * scale;
                p.y = v[e].im * scale;
                *reinterpret_cast<typename vec2<T>::type*>(dst + HO + 2 * m) = p;
                if (HO > 0) {
                    if (m == 0) dst[HO + 2 * L] = p.x;
                    if (m == L - 1) dst[HO - 1] = p.y;
                }
            }
            // the rest of the row is zeroed, so every 128-byte line of the
            // row is written whole
            if (HO == 0)
                for (long z = 2 * L + t; z < rs; z += TPL) dst[z] = 0;
            else {
                for (int z = t; z < HO - 1; z += TPL) dst[z] = 0;
                for (long z = HO + 2 * L + 1 + t; z < rs; z += TPL) dst[z] = 0;
            }
        }
        // the last exchange's LDS reads precede the next group's staging
        __syncthreads();
    }
}

// ---- z r2c pass ------------------------------------------------------------
// LOADER: functor returning the pair (x[2m], x[2m+1]) of row `row` as cx<T>.
template <typename T>
struct RealRowLoader {
    const T* __restrict__ in;
    long rs;
    long batch;
    __device__ __forceinline__ cx<T> operator()(long b, long row, int m) const {
        typename vec2<T>::type p =
            *reinterpret_cast<const typename vec2<T>::type*>(in + b * batch + row * rs + 2 * m);
        return mk<T>(p.x, p.y);
    }
};

// Rows r0 .. r1-1 are transformed.  Each work-group walks row groups
// blockIdx.x, blockIdx.x + gridDim.x, ... and loads the next group's row while
// transforming the current one.
template <typename T, int L, int E, int RW, typename LOADER, typename OP, int... Rs>
__global__ void __launch_bounds__(L / E * RW)
    k_zr2c(LOADER loader, cx<T>* __restrict__ out, long r0, long r1, int N, long hs, long out_batch,
           const cx<T>* __restrict__ tw0, const cx<T>* __restrict__ twN0, OP op) {
    constexpr int TPL = L / E;
    typedef LdsMap<T, L, RW, false> M;
    __shared__ T lre[M::SIZE];
    __shared__ T lim[M::SIZE];
    const int t0 = threadIdx.x % TPL;
    const int rl0 = threadIdx.x / TPL;
    const long ngroups = (r1 - r0 + RW - 1) / RW;
    cx<T> nxt[E];
    auto load = [&](long g) {
        const long row = r0 + g * RW + rl0;
        const bool ok = row < r1;
#pragma unroll
        for (int e = 0; e < E; e++)
            nxt[e] = ok ? loader(blockIdx.y, row, t0 + TPL * e) : mk<T>(0, 0);
    };
    if ((long)blockIdx.x < ngroups) load(blockIdx.x);
    for (long g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int t = t0, rl = rl0;
        const cx<T>* tw = tw0;
        const cx<T>* twN = twN0;
        const long row = r0 + g * RW + rl;
        const bool ok = row < r1;
        cx<T> v[E];
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = nxt[e];
        if (g + gridDim.x < ngroups) load(g + gridDim.x);
        line_fft<T, L, E, RW, false, -1, Rs...>(v, lre, lim, tw, t, rl);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int m = t + TPL * e;
            lre[M::idx(m, rl)] = v[e].re;
            lim[M::idx(m, rl)] = v[e].im;
        }
        __syncthreads();
        if (ok) {
            const int gx = (int)(row / N), gy = (int)(row % N);
            cx<T>* dst = out + blockIdx.y * out_batch + row * hs;
            const T h = (T)0.5;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = t + TPL * e;
                const int km = (k == 0) ? 0 : L - k;
                cx<T> zk = mk<T>(lre[M::idx(k, rl)], lim[M::idx(k, rl)]);
                cx<T> zm = conjg(mk<T>(lre[M::idx(km, rl)], lim[M::idx(km, rl)]));
                cx<T> ev = cscale(zk + zm, h);
                cx<T> od = cscale(zk - zm, h);           // = i * O[k]
                cx<T> w = twN[k];                        // exp(-2 pi i k / N)
                cx<T> xk = ev + mul_i<-1>(cmul(w, od));  // E + w * O
                st(dst + k, op.post(xk, gx, gy, k));
            }
            if (t == 0) {
                cx<T> z0 = mk<T>(lre[M::idx(0, rl)], lim[M::idx(0, rl)]);
                st(dst + L, op.post(mk<T>(z0.re - z0.im, (T)0), gx, gy, L));
            }
            // zeros in the row padding: whole 128-byte lines are written
            for (long k = L + 1 + t; k < hs; k += TPL) st(dst + k, mk<T>(0, 0));
        }
        // the LDS reads above precede the next group's exchanges
        __syncthreads();
    }
}

}  // namespace mitos
