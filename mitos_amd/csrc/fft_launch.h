// Host-side launch layer for the FFT passes: size dispatch, twiddle tables,
// one launcher per pass kind and the composed 3D transforms.
#pragma once

#include <map>
#include <mutex>
#include <type_traits>

#include "fft_passes.h"

namespace mitos {

// Grid sizes with compiled FFT passes (line length N for the x/y passes, N/2
// for the z passes): powers of two from 8 to 1024 and 3*2^k from 24 to 1536.
template <typename F>
int with_size(int N, F&& f) {
    switch (N) {
        case 8: return f.template operator()<8>();
        case 16: return f.template operator()<16>();
        case 24: return f.template operator()<24>();
        case 32: return f.template operator()<32>();
        case 48: return f.template operator()<48>();
        case 64: return f.template operator()<64>();
        case 96: return f.template operator()<96>();
        case 128: return f.template operator()<128>();
        case 192: return f.template operator()<192>();
        case 256: return f.template operator()<256>();
        case 384: return f.template operator()<384>();
        case 512: return f.template operator()<512>();
        case 768: return f.template operator()<768>();
        case 1024: return f.template operator()<1024>();
        case 1536: return f.template operator()<1536>();
        default: return MITOS_ERR_UNSUPPORTED;
    }
}

inline bool size_supported(int N) {
    switch (N) {
        case 8: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192:
        case 256: case 384: case 512: case 768: case 1024: case 1536:
            return true;
        default:
            return false;
    }
}

// Device twiddle table tw[k] = exp(-2 pi i k / L), k < L, one per (device, L, type).
template <typename T>
const cx<T>* twiddles(int L) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, cx<T>*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair(dev, L);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    cx<T>* h = new cx<T>[L];
    for (int k = 0; k < L; k++) {
        // reduce by symmetry so the quarter-period values are exact
        long n = k, d = L;
        double c, s;
        if (4 * n == d) { c = 0; s = 1; }
        else if (2 * n == d) { c = -1; s = 0; }
        else if (4 * n == 3 * d) { c = 0; s = -1; }
        else if (n == 0) { c = 1; s = 0; }
        else {
            const double a = 2.0 * M_PI * (double)n / (double)d;
            c = cos(a);
            s = sin(a);
        }
        h[k] = mk<T>((T)c, (T)(-s));
    }
    cx<T>* d = nullptr;
    if (hipMalloc(&d, sizeof(cx<T>) * L) != hipSuccess) {
        delete[] h;
        return nullptr;
    }
    if (copy_sync(d, h, sizeof(cx<T>) * L, hipMemcpyHostToDevice) != hipSuccess) {
        delete[] h;
        hipFree(d);
        return nullptr;
    }
    delete[] h;
    cache[key] = d;
    return d;
}

inline int launch_status() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "mitos_hip: kernel launch failed: %s\n", hipGetErrorString(e));
        return MITOS_ERR_DEVICE;
    }
    return MITOS_OK;
}

// ---- strided pass ----------------------------------------------------------
// k_pass_strided addresses a thread's elements by a 32-bit byte offset from a
// uniform base: TPL line positions of stride pstride must span under 4 GiB
template <typename T, int L, int AXIS>
inline bool strided_offsets_fit(int tpl, long hs) {
    const long pstride = AXIS == 0 ? (long)L * hs : hs;
    return (long)tpl * pstride * (long)sizeof(cx<T>) < (1L << 32);
}

// outer indices per work-group of a strided pass: 2 consecutive y rows for the
// x lines (measured: the rows of an x pass stream faster from one work-group)
template <int L, int AXIS>
struct OuterPerBlock {
    static constexpr int value = (AXIS == 0 && L >= 64) ? 2 : 1;
};

template <typename T, int L, int C, int AXIS, int DIR, bool SPLIT, typename OP, typename RAD>
struct StridedK;
template <typename T, int L, int C, int AXIS, int DIR, bool SPLIT, typename OP, int... Rs>
struct StridedK<T, L, C, AXIS, DIR, SPLIT, OP, Radices<Rs...>> {
    static void go(dim3 g, dim3 b, hipStream_t s, const cx<T>* in, cx<T>* out, long hs,
                   int ncols, long bstride, const cx<T>* tw, OP op) {
        hipLaunchKernelGGL((k_pass_strided<T, L, LineCfg<L>::E, C, AXIS, DIR, SPLIT,
                                           OuterPerBlock<L, AXIS>::value, OP, Rs...>),
                           g, b, 0, s, in, out, hs, ncols, bstride, tw, op, 0);
    }
};

template <typename T, int L, int AXIS, int DIR, typename OP>
int strided_pass(const cx<T>* in, cx<T>* out, long hs, int nbatch, long bstride, OP op,
                 hipStream_t s) {
    constexpr int C = ColTile<T>::C;
    constexpr int E = LineCfg<L>::E;
    constexpr int TPL = L / E;
    constexpr bool SPLIT = 2 * LdsMap<T, L, C, true>::SIZE * sizeof(T) > 81920;
    const int ncols = L / 2 + 1;
    if (!strided_offsets_fit<T, L, AXIS>(TPL, hs)) return MITOS_ERR_UNSUPPORTED;
    const cx<T>* tw = twiddles<T>(L);
    if (!tw) return MITOS_ERR_ALLOC;
    dim3 grid((stored_cols<T, C>(ncols, hs) + C - 1) / C, L / OuterPerBlock<L, AXIS>::value,
              nbatch);
    StridedK<T, L, C, AXIS, DIR, SPLIT, OP, typename LineCfg<L>::R>::go(
        grid, dim3(TPL * C), s, in, out, hs, ncols, bstride, tw, op);
    return launch_status();
}

// ---- z passes ----------------------------------------------------------------
template <int TPL>
struct RowsPerBlock {
    static constexpr int value = TPL >= 256 ? 1 : 256 / TPL;
};

template <typename T, int L, int RW, int HO, typename OP, typename RAD>
struct ZC2RK;
template <typename T, int L, int RW, int HO, typename OP, int... Rs>
struct ZC2RK<T, L, RW, HO, OP, Radices<Rs...>> {
    template <typename... A>
    static void go(dim3 g, dim3 b, hipStream_t s, A... a) {
        hipLaunchKernelGGL((k_zc2r<T, L, LineCfg<L>::E, RW, HO, OP, Rs...>), g, b, 0, s, a...);
    }
};

// HO > 0 writes rows with a periodic one-value halo (see k_zc2r)
template <typename T, int N, typename OP, int HO = 0>
int zc2r_pass(const cx<T>* in, T* out, long hs, long rs, int nbatch, long inb, long outb,
              T scale, OP op, hipStream_t s) {
    if (rs < N + (HO > 0 ? HO + 1 : 0)) return MITOS_ERR_ARG;
    constexpr int L = N / 2;
    constexpr int TPL = L / LineCfg<L>::E;
    constexpr int RW = RowsPerBlock<TPL>::value;
    const long nrows = (long)N * N;
    const cx<T>* tw = twiddles<T>(L);
    const cx<T>* twN = twiddles<T>(N);
    if (!tw || !twN) return MITOS_ERR_ALLOC;
    dim3 grid((unsigned)((nrows + RW - 1) / RW), nbatch);
    ZC2RK<T, L, RW, HO, OP, typename LineCfg<L>::R>::go(grid, dim3(TPL * RW), s, in, out, 0L,
                                                     nrows, N, hs, rs, inb, outb, tw, twN, scale,
                                                     op);
    return launch_status();
}

template <typename T, int L, int RW, typename LOADER, typename OP, typename RAD>
struct ZR2CK;
template <typename T, int L, int RW, typename LOADER, typename OP, int... Rs>
struct ZR2CK<T, L, RW, LOADER, OP, Radices<Rs...>> {
    template <typename... A>
    static void go(dim3 g, dim3 b, hipStream_t s, A... a) {
        hipLaunchKernelGGL((k_zr2c<T, L, LineCfg<L>::E, RW, LOADER, OP, Rs...>), g, b, 0, s,
                           a...);
    }
};

template <typename T, int N, typename LOADER, typename OP>
int zr2c_pass(LOADER loader, cx<T>* out, long hs, int nbatch, long outb, OP op, hipStream_t s) {
    constexpr int L = N / 2;
    constexpr int TPL = L / LineCfg<L>::E;
    constexpr int RW = RowsPerBlock<TPL>::value;
    const long nrows = (long)N * N;
    const cx<T>* tw = twiddles<T>(L);
    const cx<T>* twN = twiddles<T>(N);
    if (!tw || !twN) return MITOS_ERR_ALLOC;
    dim3 grid((unsigned)((nrows + RW - 1) / RW), nbatch);
    ZR2CK<T, L, RW, LOADER, OP, typename LineCfg<L>::R>::go(grid, dim3(TPL * RW), s, loader,
                                                             out, 0L, nrows, N, hs, outb, tw, twN,
                                                             op);
    return launch_status();
}

// ---- tunable schedules ---------------------------------------------------------
// Variants of the plain passes that the autotuner (autotune.h) times on the
// device and picks from, per (device, N, precision):
//   strided  0: LineCfg radices, 128-byte column tile
//            1: LineCfgS radices (fewer values per thread), the widest tile
//               of at most 1024 threads
//            2: 256-byte column tile when it fits (32 values per thread for
//               1024-point float lines)
//            3: LineCfg radices, 128-byte tile, two LDS exchange buffers
//               (one exchange step fewer) when they fit in 160 KiB
//   z passes 0: LineCfg radices, RowsPerBlock rows per work-group
//            1: LineCfgS radices, RowsPerBlock rows of its thread count
//            2: LineCfg radices, twice the rows when it fits
//            3: LineCfg radices, one row per work-group when a row fills a
//               wave (no barrier waits on other rows)
constexpr int kFftVariants = 4;

// 32 values per thread for the 256-byte tile of 1024-point float lines
template <int L>
struct LineCfgW {
    static constexpr int E = LineCfg<L>::E;
    typedef typename LineCfg<L>::R R;
};
template <>
struct LineCfgW<1024> {
    static constexpr int E = 32;
    typedef Radices<16, 16, 4> R;
};

template <typename T, int L, int V>
struct StridedSched {
    typedef typename std::conditional<
        V == 1, LineCfgS<L>,
        typename std::conditional<V == 2 && sizeof(T) == 4, LineCfgW<L>, LineCfg<L>>::type>::type
        CFG;
    static constexpr int TPL = L / CFG::E;
    static constexpr int CB = (TPL * ColTile<T>::C <= 1024) ? ColTile<T>::C : 1024 / TPL;
    // the wider tile only while one exchange buffer fits in the LDS
    static constexpr int C2 = 2 * ColTile<T>::C;
    static constexpr bool WIDE = V == 2 && TPL * C2 <= 1024 &&
                                 (long)L * (C2 + 1) * (long)sizeof(T) <= 163840;
    static constexpr int C = WIDE ? C2 : CB;
};

template <typename T, int L, int C, int AXIS, int DIR, bool SPLIT, typename OP, typename CFG,
          typename RAD = typename CFG::R>
struct StridedKV;
template <typename T, int L, int C, int AXIS, int DIR, bool SPLIT, typename OP, typename CFG,
          int... Rs>
struct StridedKV<T, L, C, AXIS, DIR, SPLIT, OP, CFG, Radices<Rs...>> {
    static void go(dim3 g, dim3 b, hipStream_t s, const cx<T>* in, cx<T>* out, long hs,
                   int ncols, long bstride, const cx<T>* tw, OP op, int o0) {
        hipLaunchKernelGGL((k_pass_strided<T, L, CFG::E, C, AXIS, DIR, SPLIT,
                                           OuterPerBlock<L, AXIS>::value, OP, Rs...>),
                           g, b, 0, s, in, out, hs, ncols, bstride, tw, op, o0);
    }
};

// outer indices o0 .. o0+no-1 only when no > 0
template <typename T, int L, int AXIS, int DIR, int V, typename OP>
int strided_pass_v(const cx<T>* in, cx<T>* out, long hs, int nbatch, long bstride, OP op,
                   hipStream_t s, int o0 = 0, int no = 0) {
    if (no <= 0) {
        o0 = 0;
        no = L;
    }
    constexpr int KO = OuterPerBlock<L, AXIS>::value;
    if (o0 < 0 || o0 + no > L || o0 % KO || no % KO) return MITOS_ERR_ARG;
    typedef StridedSched<T, L, V> S;
    constexpr int C = S::C;
    constexpr long LDS2 = 2 * (long)LdsMap<T, L, C, true>::SIZE * (long)sizeof(T);
    constexpr bool SPLIT = V == 3 ? LDS2 > 163840 : LDS2 > 81920;
    const int ncols = L / 2 + 1;
    if (!strided_offsets_fit<T, L, AXIS>(S::TPL, hs)) return MITOS_ERR_UNSUPPORTED;
    const cx<T>* tw = twiddles<T>(L);
    if (!tw) return MITOS_ERR_ALLOC;
    dim3 grid((stored_cols<T, C>(ncols, hs) + C - 1) / C, no / KO, nbatch);
    StridedKV<T, L, C, AXIS, DIR, SPLIT, OP, typename S::CFG>::go(
        grid, dim3(S::TPL * C), s, in, out, hs, ncols, bstride, tw, op, o0);
    return launch_status();
}

// strided pass with a run-time schedule variant
template <typename T, int L, int AXIS, int DIR, typename OP>
int strided_pass_tuned(int variant, const cx<T>* in, cx<T>* out, long hs, int nbatch,
                       long bstride, OP op, hipStream_t s, int o0 = 0, int no = 0) {
    switch (variant) {
        case 1:
            return strided_pass_v<T, L, AXIS, DIR, 1>(in, out, hs, nbatch, bstride, op, s, o0, no);
        case 2:
            return strided_pass_v<T, L, AXIS, DIR, 2>(in, out, hs, nbatch, bstride, op, s, o0, no);
        case 3:
            return strided_pass_v<T, L, AXIS, DIR, 3>(in, out, hs, nbatch, bstride, op, s, o0, no);
        default:
            return strided_pass_v<T, L, AXIS, DIR, 0>(in, out, hs, nbatch, bstride, op, s, o0, no);
    }
}

// work-groups of a z pass: each walks several row groups (about 8192 waves
// in all, enough to fill the device), at most one per group
inline unsigned z_groups(long nrows, int rw, int tpl) {
    const long ngroups = (nrows + rw - 1) / rw;
    const long waves = ((long)tpl * rw + 63) / 64;
    const long cap = 8192 / waves > 0 ? 8192 / waves : 1;
    return (unsigned)(ngroups < cap ? ngroups : cap);
}

template <typename T, int L, int V>
struct ZSched {
    typedef typename std::conditional<V == 1, LineCfgS<L>, LineCfg<L>>::type CFG;
    static constexpr int TPL = L / CFG::E;
    static constexpr int RW0 = V == 3 ? (TPL >= 64 ? 1 : RowsPerBlock<TPL>::value)
                                      : RowsPerBlock<TPL>::value * (V == 2 ? 2 : 1);
    static constexpr int RW = (TPL * RW0 <= 1024) ? RW0 : RowsPerBlock<TPL>::value;
};

template <typename T, int L, int RW, int HO, typename OP, typename CFG,
          typename RAD = typename CFG::R>
struct ZC2RKV;
template <typename T, int L, int RW, int HO, typename OP, typename CFG, int... Rs>
struct ZC2RKV<T, L, RW, HO, OP, CFG, Radices<Rs...>> {
    template <typename... A>
    static void go(dim3 g, dim3 b, hipStream_t s, A... a) {
        hipLaunchKernelGGL((k_zc2r<T, L, CFG::E, RW, HO, OP, Rs...>), g, b, 0, s, a...);
    }
};

// rows r0 .. r1-1 only when r1 > 0
template <typename T, int N, int V, int HO, typename OP>
int zc2r_pass_v(const cx<T>* in, T* out, long hs, long rs, int nbatch, long inb, long outb,
                T scale, OP op, hipStream_t s, long r0 = 0, long r1 = 0) {
    if (rs < N + (HO > 0 ? HO + 1 : 0)) return MITOS_ERR_ARG;
    constexpr int L = N / 2;
    typedef ZSched<T, L, V> S;
    if (r1 <= 0) {
        r0 = 0;
        r1 = (long)N * N;
    }
    if (r0 < 0 || r1 > (long)N * N || r0 >= r1) return MITOS_ERR_ARG;
    const cx<T>* tw = twiddles<T>(L);
    const cx<T>* twN = twiddles<T>(N);
    if (!tw || !twN) return MITOS_ERR_ALLOC;
    dim3 grid(z_groups(r1 - r0, S::RW, S::TPL), nbatch);
    ZC2RKV<T, L, S::RW, HO, OP, typename S::CFG>::go(grid, dim3(S::TPL * S::RW), s, in, out, r0,
                                                     r1, N, hs, rs, inb, outb, tw, twN, scale,
                                                     op);
    return launch_status();
}

template <typename T, int N, int HO, typename OP>
int zc2r_pass_tuned(int variant, const cx<T>* in, T* out, long hs, long rs, int nbatch, long inb,
                    long outb, T scale, OP op, hipStream_t s, long r0 = 0, long r1 = 0) {
    switch (variant) {
        case 1:
            return zc2r_pass_v<T, N, 1, HO>(in, out, hs, rs, nbatch, inb, outb, scale, op, s, r0,
                                            r1);
        case 2:
            return zc2r_pass_v<T, N, 2, HO>(in, out, hs, rs, nbatch, inb, outb, scale, op, s, r0,
                                            r1);
        case 3:
            return zc2r_pass_v<T, N, 3, HO>(in, out, hs, rs, nbatch, inb, outb, scale, op, s, r0,
                                            r1);
        default:
            return zc2r_pass_v<T, N, 0, HO>(in, out, hs, rs, nbatch, inb, outb, scale, op, s, r0,
                                            r1);
    }
}

template <typename T, int L, int RW, typename LOADER, typename OP, typename CFG,
          typename RAD = typename CFG::R>
struct ZR2CKV;
template <typename T, int L, int RW, typename LOADER, typename OP, typename CFG, int... Rs>
struct ZR2CKV<T, L, RW, LOADER, OP, CFG, Radices<Rs...>> {
    template <typename... A>
    static void go(dim3 g, dim3 b, hipStream_t s, A... a) {
        hipLaunchKernelGGL((k_zr2c<T, L, CFG::E, RW, LOADER, OP, Rs...>), g, b, 0, s, a...);
    }
};

// rows r0 .. r1-1 only when r1 > 0
template <typename T, int N, int V, typename LOADER, typename OP>
int zr2c_pass_v(LOADER loader, cx<T>* out, long hs, int nbatch, long outb, OP op,
                hipStream_t s, long r0 = 0, long r1 = 0) {
    constexpr int L = N / 2;
    typedef ZSched<T, L, V> S;
    if (r1 <= 0) {
        r0 = 0;
        r1 = (long)N * N;
    }
    if (r0 < 0 || r1 > (long)N * N || r0 >= r1) return MITOS_ERR_ARG;
    const cx<T>* tw = twiddles<T>(L);
    const cx<T>* twN = twiddles<T>(N);
    if (!tw || !twN) return MITOS_ERR_ALLOC;
    dim3 grid(z_groups(r1 - r0, S::RW, S::TPL), nbatch);
    ZR2CKV<T, L, S::RW, LOADER, OP, typename S::CFG>::go(grid, dim3(S::TPL * S::RW), s, loader,
                                                         out, r0, r1, N, hs, outb, tw, twN, op);
    return launch_status();
}

template <typename T, int N, typename LOADER, typename OP>
int zr2c_pass_tuned(int variant, LOADER loader, cx<T>* out, long hs, int nbatch, long outb,
                    OP op, hipStream_t s, long r0 = 0, long r1 = 0) {
    switch (variant) {
        case 1: return zr2c_pass_v<T, N, 1>(loader, out, hs, nbatch, outb, op, s, r0, r1);
        case 2: return zr2c_pass_v<T, N, 2>(loader, out, hs, nbatch, outb, op, s, r0, r1);
        case 3: return zr2c_pass_v<T, N, 3>(loader, out, hs, nbatch, outb, op, s, r0, r1);
        default: return zr2c_pass_v<T, N, 0>(loader, out, hs, nbatch, outb, op, s, r0, r1);
    }
}

// ---- composed 3D transforms ---------------------------------------------------
// c2r: x (in -> ws), y (ws -> ws), z (ws -> out).  The input is left unchanged.
template <typename T, typename OPX, typename OPY, typename OPZ>
int c2r_3d(const cx<T>* in, cx<T>* ws, T* out, const GridLayout& g, int nb, long cb, long rb,
           T scale, OPX ox, OPY oy, OPZ oz, hipStream_t s) {
    return with_size(g.N, [&]<int NN>() -> int {
        MITOS_TRY((strided_pass<T, NN, 0, 1>(in, ws, g.hs, nb, cb, ox, s)));
        MITOS_TRY((strided_pass<T, NN, 1, 1>(ws, ws, g.hs, nb, cb, oy, s)));
        return zc2r_pass<T, NN>(ws, out, g.hs, g.rs, nb, cb, rb, scale, oz, s);
    });
}

// r2c: z (loader -> out), y (out -> out), x (out -> out).
template <typename T, typename LOADER, typename OPZ, typename OPY, typename OPX>
int r2c_3d(LOADER loader, cx<T>* out, const GridLayout& g, int nb, long cb, OPZ oz, OPY oy,
           OPX ox, hipStream_t s) {
    return with_size(g.N, [&]<int NN>() -> int {
        MITOS_TRY((zr2c_pass<T, NN>(loader, out, g.hs, nb, cb, oz, s)));
        MITOS_TRY((strided_pass<T, NN, 1, -1>(out, out, g.hs, nb, cb, oy, s)));
        return strided_pass<T, NN, 0, -1>(out, out, g.hs, nb, cb, ox, s);
    });
}

}  // namespace mitos
