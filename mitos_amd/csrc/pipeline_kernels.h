// Fused kernels of the initial-condition pipeline.
//
// Per realization the pipeline keeps every field in HBM and replaces the
// reference's 19 separate 3D FFTs plus elementwise passes by 11 transforms'
// worth of fused passes:
//   noise+spectrum -> phi1(k) = -delta(k)/k^2
//   x-pass (1 in, 3 out), y-pass (3 in, 6 out): the six Hessian components in
//     (x, y, k_z) space, factored so each FFT line is read once
//   z-pass: six c2r rows -> delta2 = sum of Hessian minors -> r2c row
//   y-pass forward, x-pass forward + Poisson + 2LPT combine + TSC window +
//     first inverse pass of the displacement fields (1 in + phi1, 2 out)
//   y-pass (2 in, 3 out), z c2r x3 -> displacement grids
//   TSC displacement of the lattice, fixed-point CIC, r2c with the integer
//     grid converted on load, P(k) with the CIC window and normalisation.
#pragma once

#include "fft_passes.h"
#include "kspace.h"
#include "particles.h"

namespace mitos {

// ---- multi-output strided pass ------------------------------------------------
// Output j = FFT_DIR( m_j(k_axis) * in[src_j] ) where the multiplier code is
// 0: 1, 1: i k, 2: -k^2, 3: i k S (k = signed frequency of the line position
// times dk, S = smooth[line position]).
template <typename T, int NIN, int NOUT>
struct MultiPtrs {
    const cx<T>* in[NIN];
    cx<T>* out[NOUT];
    const T* smooth = nullptr;  // per line position, for multiplier code 3
};

template <int... CODES>
struct OutCodes {
    static constexpr int n = sizeof...(CODES);
    static constexpr int code[sizeof...(CODES)] = {CODES...};
    static constexpr int src(int j) { return code[j] >> 2; }
    static constexpr int mul(int j) { return code[j] & 3; }
};
#define MITOS_OUT(src, mul) ((src) * 4 + (mul))

template <typename T>
__device__ __forceinline__ cx<T> apply_mul(int code, cx<T> v, T k) {
    if (code == 1 || code == 3) return mk<T>(-v.im * k, v.re * k);  // i k v
    if (code == 2) {
        const T m = -(k * k);
        return mk<T>(v.re * m, v.im * m);
    }
    return v;
}

// Work-group (bx, by) handles column tile bx for the KO outer indices
// by*KO .. by*KO + KO-1 (consecutive y rows of an x pass stream faster from
// one work-group).
template <typename T, int L, int E, int C, int AXIS, int DIR, bool SPLIT, int NIN,
          typename CODES, int KO, int... Rs>
__global__ void __launch_bounds__(L / E * C)
    k_multi_strided(MultiPtrs<T, NIN, CODES::n> P, long hs, int ncols, double dk,
                    const cx<T>* __restrict__ tw) {
    constexpr int TPL = L / E;
    constexpr int NOUT = CODES::n;
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
    // loads: uniform base per element plus one 32-bit per-thread offset
    const unsigned toff = (unsigned)(((long)t * pstride + c) * (long)sizeof(cx<T>));
#pragma unroll 1
    for (int i = 0; i < KO; i++) {
        const int o = (int)by * KO + i;
        const long base = (long)o * ostride + col;
        const long ubase = (long)o * ostride + (long)bx * C;
        cx<T> v[E], u[E];
#pragma unroll
        for (int j = 0; j < NOUT; j++) {
            // per-output copies: LDS addresses and twiddles shared by all
            // outputs would otherwise stay live across the whole kernel
            int tl = t, cl = c;
            const cx<T>* twl = tw;
            asm volatile("" : "+v"(tl), "+v"(cl), "+s"(twl));
            if (j == 0 || CODES::src(j) != CODES::src(j > 0 ? j - 1 : 0)) {
                const char* ib = reinterpret_cast<const char*>(P.in[CODES::src(j)] + ubase);
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const char* eb = ib + (long)TPL * e * pstride * (long)sizeof(cx<T>);
                    v[e] = ok ? ld(reinterpret_cast<const cx<T>*>(eb + toff)) : mk<T>(0, 0);
                }
            }
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int pos = tl + TPL * e;
                T k = (T)(freq_index(pos, L) * dk);
                if (CODES::mul(j) == 3) k *= P.smooth[pos];
                u[e] = apply_mul(CODES::mul(j), v[e], k);
            }
            line_fft<T, L, E, C, true, DIR, Rs...>(u, lds, SPLIT ? lds : lds + M::SIZE, twl, tl,
                                                   cl);
            // padding columns of the last tile hold zeros (zero input, linear map)
            if (wok) {
                cx<T>* out = P.out[j];
#pragma unroll
                for (int e = 0; e < E; e++)
                    st(out + base + (long)(tl + TPL * e) * pstride, u[e]);
            }
        }
    }
}


// ---- z pass: six Hessian rows -> delta2 row (reals) ---------------------------- ----------------------------
// Inputs (rows of hs complex, (x, y, k_z) space): B[0] = Y(A0), B[1] = Y(i ky A0),
// B[2] = Y(-ky^2 A0), B[3] = Y(A1), B[4] = Y(i ky A1), B[5] = Y(A2).
// Hessian rows: xx = B5, yy = B2, zz = -kz^2 B0, xy = B4, yz = i kz B1,
// xz = i kz B3.  delta2 follows src/2lpt.c:111-112 term by term.
// The delta2 row is written as N reals (row stride rs); its r2c runs as a
// separate z pass (a fused r2c here doubles the live registers).
template <typename T, int L, int E, int RW, int... Rs>
__global__ void __launch_bounds__(L / E * RW)
    k_hess_z(MultiPtrs<T, 6, 1> P, long nrows, long hs, double dk, T scale,
             const cx<T>* __restrict__ tw0, const cx<T>* __restrict__ twN0, T* __restrict__ real,
             long rs) {
    constexpr int TPL = L / E;
    typedef LdsMap<T, L, RW, false> M;
    __shared__ T lre[M::SIZE];
    __shared__ T lim[M::SIZE];
    const int t0 = threadIdx.x % TPL;
    const int rl0 = threadIdx.x / TPL;
    const long row = (long)blockIdx.x * RW + rl0;
    const bool ok = row < nrows;
    // xx -> A; yy -> B = xx + yy, A = xx * yy; zz -> A = zz * B + A; then
    // A -= h^2 for xy, yz, xz: two accumulator arrays besides the transform
    cx<T> A[E], B[E], cur[E];
    // component order: xx, yy, zz, xy, yz, xz; the next component's row is
    // loaded while the current one is transformed
    auto src_of = [](int q) {
        return (q == 0) ? 5 : (q == 1) ? 2 : (q == 2) ? 0 : (q == 3) ? 4 : (q == 4) ? 1 : 3;
    };
    cx<T> nxt[E + 1];
    auto load_row = [&](int q) {
        const cx<T>* src = P.in[src_of(q)] + row * hs;
#pragma unroll
        for (int e = 0; e < E; e++) nxt[e] = ld(src + t0 + TPL * e);
        nxt[E] = t0 == 0 ? ld(src + L) : mk<T>(0, 0);
    };
    if (ok) load_row(0);
#pragma unroll 1
    for (int q = 0; q < 6; q++) {
        const int t = t0, rl = rl0;
        const cx<T>* tw = tw0;
        const cx<T>* twN = twN0;
        const int mul_q = (q == 2) ? 2 : (q >= 4) ? 1 : 0;
        // stage the row, with its k_z multiplier, in LDS
        if (ok) {
#pragma unroll
            for (int e = 0; e <= E; e++) {
                const int k = (e < E) ? t + TPL * e : L;
                if (e == E && t != 0) break;
                cx<T> a = apply_mul(mul_q, nxt[e], (T)(k * dk));
                lre[M::idx(k, rl)] = a.re;
                lim[M::idx(k, rl)] = a.im;
            }
            if (q + 1 < 6) load_row(q + 1);
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = t + TPL * e;
            cx<T> xk = mk<T>(lre[M::idx(k, rl)], lim[M::idx(k, rl)]);
            cx<T> xm = mk<T>(lre[M::idx(L - k, rl)], lim[M::idx(L - k, rl)]);
            if (k == 0) {
                xk.im = 0;
                xm.im = 0;
            }
            cx<T> a = xk + conjg(xm);
            cx<T> b = xk - conjg(xm);
            cx<T> w = twN[k];
            w.im = -w.im;
            cur[e] = a + mul_i<1>(cmul(w, b));
        }
        __syncthreads();
        line_fft<T, L, E, RW, false, 1, Rs...>(cur, lre, lim, tw, t, rl);
#pragma unroll
        for (int e = 0; e < E; e++) {
            // real values x[2m] = re, x[2m+1] = im
            const cx<T> h = mk<T>(cur[e].re * scale, cur[e].im * scale);
            if (q == 0) {
                A[e] = h;
            } else if (q == 1) {
                B[e] = mk<T>(A[e].re + h.re, A[e].im + h.im);
                A[e] = mk<T>(A[e].re * h.re, A[e].im * h.im);
            } else if (q == 2) {
                // d_dzz*d_dyy + d_dzz*d_dxx + d_dyy*d_dxx = d_dzz*(d_dxx+d_dyy) + d_dxx*d_dyy
                A[e].re = h.re * B[e].re + A[e].re;
                A[e].im = h.im * B[e].im + A[e].im;
            } else {
                // - d_dxy^2 - d_dyz^2 - d_dxz^2
                A[e].re = A[e].re - h.re * h.re;
                A[e].im = A[e].im - h.im * h.im;
            }
        }
    }
    if (!ok) return;
    T* dst = real + row * rs;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int m = t0 + TPL * e;
        typename vec2<T>::type p;
        p.x = A[e].re;
        p.y = A[e].im;
        *reinterpret_cast<typename vec2<T>::type*>(dst + 2 * m) = p;
    }
}

// ---- x pass: delta2 forward + Poisson + 2LPT combine + window + 2 inverse ----
struct CombineParams {
    double dk;
    double sD;      // forward normalisation L^3/N^3 of delta2
    double f1, f2;  // 2LPT factors
    const double* win;  // 1/sinc^3 per axis (TSC window), or nullptr
    // lattice TSC stencil per axis, S[n] = 3/4 + cos(2 pi n/N)/4, or nullptr:
    // C0 (the y and z displacements) takes S(kz), C1 (x) S(kx) S(ky) S(kz)
    const double* smooth;
};

// two 1024-thread work-groups per CU need at most 64 registers per thread
// (float lines up to 1024 points; the others keep their registers)
template <typename T, int L, int E, int C, bool SPLIT, int... Rs>
__global__ void __launch_bounds__(L / E * C) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && L <= 1024 ? 8 : 1)))
    k_x_combine(const cx<T>* D, const cx<T>* __restrict__ K, cx<T>* C0, cx<T>* C1, long hs,
                int ncols, CombineParams cp, const cx<T>* __restrict__ tw) {
#pragma clang fp contract(off)
    constexpr int TPL = L / E;
    typedef LdsMap<T, L, C, true> M;
    __shared__ T lds[(SPLIT ? 1 : 2) * M::SIZE];
    const int c = threadIdx.x % C;
    const int t = threadIdx.x / C;
    unsigned bx, by;
    xcd_block2(bx, by);
    const int col = bx * C + c;
    const int y = by;
    const bool ok = col < ncols;
    const bool wok = col < stored_cols<T, C>(ncols, hs);
    const long pstride = (long)L * hs;
    const long base = (long)y * hs + col;
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; e++)
        v[e] = ok ? ld(D + base + (long)(t + TPL * e) * pstride) : mk<T>(0, 0);
    line_fft<T, L, E, C, true, -1, Rs...>(v, lds, SPLIT ? lds : lds + M::SIZE, tw, t, c);
    // per-mode arithmetic in the grid's precision (double for FP64)
    typedef T W;
    const W ky = (W)(freq_index(y, L) * cp.dk);
    const W kz = (W)((ok ? col : 0) * cp.dk);
    const W wyz = cp.win ? (W)(cp.win[y] * cp.win[ok ? col : 0]) : (W)1;
    const W kyz2 = ky * ky + kz * kz;
    const W f1 = (W)cp.f1, f2sD = (W)cp.f2 * (W)cp.sD;
    cx<T> u[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int x = t + TPL * e;
        const W kx = (W)(freq_index(x, L) * cp.dk);
        const W k2 = kx * kx + kyz2;
        const cx<T> phi1 = ok ? ld(K + base + (long)x * pstride) : mk<T>(0, 0);
        // phi2 = inv_poisson(delta2(k)) (include/fft_kernels.h:96-99), DC x 1
        const W ip = (k2 > 0) ? (W)-1 / k2 : (W)1;
        W pre, pim;
        if constexpr (sizeof(T) == 8) {
            pre = phi1.re * cp.f1 + cp.f2 * (v[e].re * cp.sD * ip);
            pim = phi1.im * cp.f1 + cp.f2 * (v[e].im * cp.sD * ip);
        } else {
            pre = phi1.re * f1 + f2sD * (v[e].re * ip);
            pim = phi1.im * f1 + f2sD * (v[e].im * ip);
        }
        if (cp.win && k2 > 0) {
            const W iw = (W)cp.win[x] * wyz;  // 1/W_TSC(k), include/fft_kernels.h:198-219
            pre *= iw;
            pim *= iw;
        }
        if (cp.smooth) {
            const W sz = (W)cp.smooth[ok ? col : 0];
            pre *= sz;
            pim *= sz;
            // x displacement: i kx S(kx) S(ky) times the value
            const W kxs = kx * (W)(cp.smooth[x] * cp.smooth[y]);
            u[e] = mk<T>((T)(-pim * kxs), (T)(pre * kxs));
        }
        v[e] = mk<T>((T)pre, (T)pim);
    }
    if (!cp.smooth) {
#pragma unroll
        for (int e = 0; e < E; e++)
            u[e] = apply_mul(1, v[e], (T)(freq_index(t + TPL * e, L) * cp.dk));
    }
    // padding columns of the last tile hold zeros (zero input, linear map).
    // Each transform recomputes its LDS addresses and twiddles: shared by the
    // three they would stay live across the kernel.
    {
        int tl = t, cl = c;
        const cx<T>* twl = tw;
        asm volatile("" : "+v"(tl), "+v"(cl), "+s"(twl));
        line_fft<T, L, E, C, true, 1, Rs...>(v, lds, SPLIT ? lds : lds + M::SIZE, twl, tl, cl);
    }
    if (wok) {
#pragma unroll
        for (int e = 0; e < E; e++) st(C0 + base + (long)(t + TPL * e) * pstride, v[e]);
    }
    {
        int tl = t, cl = c;
        const cx<T>* twl = tw;
        asm volatile("" : "+v"(tl), "+v"(cl), "+s"(twl));
        line_fft<T, L, E, C, true, 1, Rs...>(u, lds, SPLIT ? lds : lds + M::SIZE, twl, tl, cl);
    }
    if (wok) {
#pragma unroll
        for (int e = 0; e < E; e++) st(C1 + base + (long)(t + TPL * e) * pstride, u[e]);
    }
}

// ---- loaders / operators for the remaining passes -------------------------------
// z c2r of the z displacement: multiply by i k_z on load
struct OpIkz {
    double dk;
    template <typename T>
    __device__ __forceinline__ cx<T> pre(cx<T> v, int, int, int z) const {
        return apply_mul(1, v, (T)(z * dk));
    }
    template <typename T>
    __device__ __forceinline__ cx<T> post(cx<T> v, int, int, int) const {
        return v;
    }
};

// density r2c: convert the fixed-point CIC grid to delta on load
template <typename T>
struct CicLoader {
    const unsigned long long* __restrict__ grid;
    long N;
    double norm;  // 2^-40 * N^3 / npart
    __device__ __forceinline__ cx<T> operator()(long, long row, int m) const {
        const ulonglong2 q = *reinterpret_cast<const ulonglong2*>(grid + row * N + 2 * m);
        return mk<T>((T)((double)q.x * norm - 1.0), (T)((double)q.y * norm - 1.0));
    }
};

// P(k) value transform: forward normalisation and the CIC window (order 2)
// from the separable table (analyse/src/mitos_read.c:414-417)
struct PkCicTransform {
    double s;            // L^3/N^3
    const double* win;   // 1/sinc^2 per axis
    int N;
    // factor of a row (x, y): s * w[x] * w[y]; the mode value is v * row * w[z]
    __device__ __forceinline__ double row(int x, int y) const { return s * win[x] * win[y]; }
    __device__ __forceinline__ double col(int z) const { return win[z]; }
};

}  // namespace mitos
