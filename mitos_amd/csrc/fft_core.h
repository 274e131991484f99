// Line FFTs for gfx950: mixed-radix Stockham passes held in VGPRs, with the
// data exchange between passes staged in LDS.
//
// A work-group transforms C independent lines of length L.  Each line is
// served by TPL = L/E threads; thread t of a line holds E complex values, the
// ones at positions t + TPL*e (e = 0..E-1), both before and after the
// transform ("register order").  This is exactly the layout a coalesced load
// of a column tile (strided passes) or of a contiguous row (z passes) gives,
// so global loads and stores need no LDS round trip.
//
// Stage s of radix R (R divides E) has L/R butterflies; thread t owns the
// butterflies j = t + TPL*b (b = 0..E/R-1), reading positions j + r*L/R and
// writing positions (j/Ns)*Ns*R + j%Ns + r*Ns, with Ns the product of the
// radices of the previous stages (Stockham autosort).  Stages exchange through
// LDS; the first stage reads and the last stage writes registers directly.
#pragma once

#include "mitos_common.h"

namespace mitos {

// ---------------------------------------------------------------------------
// in-register DFTs (DIR = -1 forward e^{-2 pi i jk/R}, +1 backward)
// ---------------------------------------------------------------------------
template <int R, int DIR, typename T>
struct DFT;

template <int DIR, typename T>
struct DFT<2, DIR, T> {
    template <int S>
    static __device__ __forceinline__ void run(cx<T>* v) {
        cx<T> a = v[0], b = v[S];
        v[0] = a + b;
        v[S] = a - b;
    }
};

template <int DIR, typename T>
struct DFT<3, DIR, T> {
    template <int S>
    static __device__ __forceinline__ void run(cx<T>* v) {
        const T h = (T)0.8660254037844386467637232;  // sqrt(3)/2
        cx<T> a = v[0], b = v[S], c = v[2 * S];
        cx<T> s1 = b + c;
        cx<T> t = mk<T>(a.re - (T)0.5 * s1.re, a.im - (T)0.5 * s1.im);
        cx<T> d = cscale(b - c, h);
        cx<T> id = mul_i<DIR>(d);
        v[0] = a + s1;
        v[S] = t + id;
        v[2 * S] = t - id;
    }
};

template <int DIR, typename T>
struct DFT<4, DIR, T> {
    template <int S>
    static __device__ __forceinline__ void run(cx<T>* v) {
        cx<T> t0 = v[0] + v[2 * S];
        cx<T> t1 = v[0] - v[2 * S];
        cx<T> t2 = v[S] + v[3 * S];
        cx<T> t3 = mul_i<DIR>(v[S] - v[3 * S]);
        v[0] = t0 + t2;
        v[2 * S] = t0 - t2;
        v[S] = t1 + t3;
        v[3 * S] = t1 - t3;
    }
};

// 8-point: two 4-point DFTs on even/odd legs, then twiddles w8^k.
template <int DIR, typename T>
struct DFT<8, DIR, T> {
    template <int S>
    static __device__ __forceinline__ void run(cx<T>* v) {
        const T c = (T)0.7071067811865475244008444;
        cx<T> e[4] = {v[0], v[2 * S], v[4 * S], v[6 * S]};
        cx<T> o[4] = {v[S], v[3 * S], v[5 * S], v[7 * S]};
        DFT<4, DIR, T>::template run<1>(e);
        DFT<4, DIR, T>::template run<1>(o);
        o[1] = cmul(o[1], mk<T>(c, DIR * c));
        o[2] = mul_i<DIR>(o[2]);
        o[3] = cmul(o[3], mk<T>(-c, DIR * c));
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k * S] = e[k] + o[k];
            v[(k + 4) * S] = e[k] - o[k];
        }
    }
};

template <int DIR, typename T>
struct DFT<16, DIR, T> {
    template <int S>
    static __device__ __forceinline__ void run(cx<T>* v) {
        const T c1 = (T)0.9238795325112867561281832;  // cos(pi/8)
        const T s1 = (T)0.3826834323650897717284600;  // sin(pi/8)
        const T r2 = (T)0.7071067811865475244008444;
        cx<T> e[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            e[k] = v[2 * k * S];
            o[k] = v[(2 * k + 1) * S];
        }
        DFT<8, DIR, T>::template run<1>(e);
        DFT<8, DIR, T>::template run<1>(o);
        o[1] = cmul(o[1], mk<T>(c1, DIR * s1));
        o[2] = cmul(o[2], mk<T>(r2, DIR * r2));
        o[3] = cmul(o[3], mk<T>(s1, DIR * c1));
        o[4] = mul_i<DIR>(o[4]);
        o[5] = cmul(o[5], mk<T>(-s1, DIR * c1));
        o[6] = cmul(o[6], mk<T>(-r2, DIR * r2));
        o[7] = cmul(o[7], mk<T>(-c1, DIR * s1));
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[k * S] = e[k] + o[k];
            v[(k + 8) * S] = e[k] - o[k];
        }
    }
};

// ---------------------------------------------------------------------------
// LDS exchange buffer: position p of line c lives at c*LINE + p*PS (+ pad)
// ---------------------------------------------------------------------------
template <typename T, int L, int C, bool COLFAST>
struct LdsMap {
    // COLFAST: lines are the fast index (strided passes, lanes walk columns)
    //   idx = p*(C+PADC) + c
    // else: positions are the fast index (z passes), one pad word after every
    // 32 positions so the radix-R scatter of an exchange (stride R in p) hits
    // distinct banks
    //   idx = c*(L + L/32 + 1) + p + p/32
    static constexpr int PADC = (C >= 4) ? 1 : 0;
    static constexpr int ROW = L + L / 32 + 1;
    static constexpr int SIZE = COLFAST ? L * (C + PADC) : C * ROW;
    static __device__ __forceinline__ int idx(int p, int c) {
        return COLFAST ? p * (C + PADC) + c : c * ROW + p + (p >> 5);
    }
};

// twiddle table access: tw[k] = exp(DIR * 2 pi i k / L) stored for DIR = -1;
// the backward twiddle is the conjugate.
template <int DIR, typename T>
__device__ __forceinline__ cx<T> twiddle(const cx<T>* __restrict__ tw, int k) {
    cx<T> w = ld(tw + k);
    if (DIR > 0) w.im = -w.im;
    return w;
}

// Recursive stage runner.
template <typename T, int L, int E, int C, bool COLFAST, int DIR, int NS, int... Rs>
struct Stages;

template <typename T, int L, int E, int C, bool COLFAST, int DIR, int NS>
struct Stages<T, L, E, C, COLFAST, DIR, NS> {
    static __device__ __forceinline__ void run(cx<T>*, T*, T*, const cx<T>*, int, int) {}
};

template <typename T, int L, int E, int C, bool COLFAST, int DIR, int NS, int R, int... Rs>
struct Stages<T, L, E, C, COLFAST, DIR, NS, R, Rs...> {
    static constexpr int TPL = L / E;
    static constexpr int NB = E / R;  // butterflies per thread in this stage
    static_assert(E % R == 0, "radix must divide elements per thread");

    static __device__ __forceinline__ void run(cx<T>* v, T* lre, T* lim,
                                               const cx<T>* __restrict__ tw, int t, int c) {
        typedef LdsMap<T, L, C, COLFAST> M;
        // v[b + r*NB] holds leg r of butterfly j = t + TPL*b
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = t + TPL * b;
            cx<T> u[R];
#pragma unroll
            for (int r = 0; r < R; r++) u[r] = v[b + r * NB];
            if (NS > 1) {
                const int jm = j % NS;
#pragma unroll
                for (int r = 1; r < R; r++)
                    u[r] = cmul(u[r], twiddle<DIR>(tw, (jm * r) * (L / (NS * R))));
            }
            DFT<R, DIR, T>::template run<1>(u);
#pragma unroll
            for (int r = 0; r < R; r++) v[b + r * NB] = u[r];
        }
        if constexpr (sizeof...(Rs) > 0) {
            // exchange through LDS into the next stage's register order
            constexpr int R2 = NextRadix<Rs...>::value;
            constexpr int NB2 = E / R2;
            if (lim != lre) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const int j = t + TPL * b;
                    const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int p = base + r * NS;
                        lre[M::idx(p, c)] = v[b + r * NB].re;
                        lim[M::idx(p, c)] = v[b + r * NB].im;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int b = 0; b < NB2; b++) {
                    const int j = t + TPL * b;
#pragma unroll
                    for (int r = 0; r < R2; r++) {
                        const int p = j + r * (L / R2);
                        v[b + r * NB2] = mk<T>(lre[M::idx(p, c)], lim[M::idx(p, c)]);
                    }
                }
                __syncthreads();
            } else {
                // one buffer: exchange real parts, then imaginary parts
                T keep[E];
#pragma unroll
                for (int half = 0; half < 2; half++) {
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        const int j = t + TPL * b;
                        const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            const int p = base + r * NS;
                            lre[M::idx(p, c)] = half ? v[b + r * NB].im : v[b + r * NB].re;
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int b = 0; b < NB2; b++) {
                        const int j = t + TPL * b;
#pragma unroll
                        for (int r = 0; r < R2; r++) {
                            const int p = j + r * (L / R2);
                            if (half) v[b + r * NB2] = mk<T>(keep[b + r * NB2], lre[M::idx(p, c)]);
                            else keep[b + r * NB2] = lre[M::idx(p, c)];
                        }
                    }
                    __syncthreads();
                }
            }
            Stages<T, L, E, C, COLFAST, DIR, NS * R, Rs...>::run(v, lre, lim, tw, t, c);
        } else {
            // Last stage: outputs of butterfly j = t+TPL*b, leg r sit at
            // position j + r*NS = t + TPL*(b + r*NB): already register order.
        }
    }

    template <int A, int... B>
    struct NextRadix {
        static constexpr int value = A;
    };
};

// Public entry: transform the E register values of every thread in place.
// lre/lim: LDS buffers of LdsMap<T,L,C,COLFAST>::SIZE elements each; passing
// lim == lre exchanges real and imaginary parts one after the other through a
// single buffer (half the LDS, two more barriers per exchange).
template <typename T, int L, int E, int C, bool COLFAST, int DIR, int... Rs>
__device__ __forceinline__ void line_fft(cx<T>* v, T* lre, T* lim, const cx<T>* __restrict__ tw,
                                         int t, int c) {
    Stages<T, L, E, C, COLFAST, DIR, 1, Rs...>::run(v, lre, lim, tw, t, c);
}

}  // namespace mitos
