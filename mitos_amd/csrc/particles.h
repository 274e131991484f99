// Particle kernels: lattice, sequential TSC displacement, fixed-point CIC.
#pragma once

#include "mitos_common.h"

#include <map>
#include <mutex>
#include <type_traits>

namespace mitos {

// fixed-point unit of the CIC deposit: weights are summed as integers of
// 2^-40 (2^-24 for float weights), so the total does not depend on the order
// of the atomic adds
#define MITOS_CIC_SHIFT 40
#define MITOS_CIC_SHIFT_F32 24

__host__ __device__ __forceinline__ double fwrap_d(double x, double L) {
    return fmod(fmod(x, L) + L, L);  // include/distributed_grid.h:27
}

// fwrap_d out of line: the far-position fallback below stays out of the
// unrolled particle loops
__host__ __device__ __noinline__ inline double fwrap_d_far(double x, double L) { return fwrap_d(x, L); }

// fwrap for -L < x < 2L without fmod: every step of fmod(fmod(x,L)+L, L) is
// an exact subtraction there, so the result has the same bits as fwrap_d.
__host__ __device__ __forceinline__ double fwrap_fast(double x, double L) {
#pragma clang fp contract(off)
    if (!(x > -L && x < 2 * L)) return fwrap_d_far(x, L);
    const double m = (x >= L) ? x - L : x;  // fmod(x, L)
    const double a = m + L;
    double r = (a >= L) ? a - L : a;         // fmod(a, L), a in [0, 2L]
    if (r >= L) r -= L;
    return r;
}

__device__ __noinline__ inline float fwrap_float_far(float x, double L) {
    return (float)fwrap_fast(x, L);
}

// Float periodic wrap of a float position: x itself on [0, Lf), Lf = (float)L
// (positions written by the displacement kernels lie there), the double wrap
// otherwise.
__device__ __forceinline__ float fwrap_float(float x, float Lf, double L) {
    return (x >= 0.0f && x < Lf) ? x : fwrap_float_far(x, L);
}

// one axis of the TSC window (src/grids_interp.c:291-296): the three taps with
// non-zero weight, in the reference's loop order.  X is in cell units.
template <typename M>
__device__ __forceinline__ void tsc_axis(M X, int& i0, M w[3]) {
#pragma clang fp contract(off)
    const int iX = (int)floor(X);
    const M d = X - iX;
    const int first = (d < (M)0.5) ? -1 : 0;  // reference loops from floor(d-1.5)
    i0 = iX + first;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const M xx = fabs(X - (M)(iX + first + t));
        w[t] = xx < (M)0.5 ? ((M)0.75 - xx * xx)
                           : (xx < (M)1.5 ? (M)0.5 * ((M)1.5 - xx) * ((M)1.5 - xx) : (M)0);
    }
}

__device__ __forceinline__ int wrapi(int i, int N) { return ((i % N) + N) % N; }
// wrap of an index known to lie in [-N, 2N)
__device__ __forceinline__ int wrap1(int i, int N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

// periodic wrap of any int index (floor division through a float reciprocal,
// then one correction each way; exact for |i| < 2^22)
__device__ __forceinline__ int wrapf(int i, int N, float invN) {
    int r = i - N * (int)floorf((float)i * invN);
    r = r < 0 ? r + N : r;
    return r >= N ? r - N : r;
}

// gridTSC_dg (src/grids_interp.c:267-309) on one rank: all taps wrap
// periodically.  The 27 taps with nonzero weight are summed in the reference's
// loop order; FMA = false keeps every product and sum separately rounded (the
// reference's arithmetic), FMA = true fuses them (FP32 variant).  Element
// offsets are 32-bit (rs * N * N < 2^32).
template <typename T, typename M, bool FMA>
__device__ __forceinline__ M tsc_gather(const T* __restrict__ box, unsigned rs, int N, float invN,
                                        M boxlen, M x, M y, M z) {
#pragma clang fp contract(off)
    const M X = x * N / boxlen;
    const M Y = y * N / boxlen;
    const M Z = z * N / boxlen;
    int i0, j0, k0;
    M wx[3], wy[3], wz[3];
    tsc_axis(X, i0, wx);
    tsc_axis(Y, j0, wy);
    tsc_axis(Z, k0, wz);
    const int ib = wrapf(i0, N, invN), jb = wrapf(j0, N, invN), kb = wrapf(k0, N, invN);
    unsigned ox[3], oy[3], oz[3];
    const unsigned plane = (unsigned)N * rs;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        ox[t] = (unsigned)wrap1(ib + t, N) * plane;
        oy[t] = (unsigned)wrap1(jb + t, N) * rs;
        oz[t] = (unsigned)wrap1(kb + t, N);
    }
    M sum = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const unsigned row = ox[a] + oy[b];
            const M wab = wx[a] * wy[b];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const M v = (M)box[row + oz[c]];
                if (FMA)
                    sum = fma(v, wab * wz[c], sum);
                else
                    sum += v * (wab * wz[c]);
            }
        }
    }
    return sum;
}

// src/particle.c:87-129: lattice position of particle p (index x*M*M+y*M+z)
__device__ __forceinline__ void lattice_pos(long long p, int M, float spacing, float& X,
                                            float& Y, float& Z) {
    const int x = (int)(p / ((long long)M * M));
    const int y = (int)((p / M) % M);
    const int z = (int)(p % M);
    X = x * spacing;
    Y = y * spacing;
    Z = z * spacing;
}

template <typename T>
__global__ void k_lattice(int M, float spacing, long long first_id, float* X, float* Y, float* Z,
                          long long* ids) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)M * M * M;
    if (p >= n) return;
    float x, y, z;
    lattice_pos(p, M, spacing, x, y, z);
    X[p] = x;
    Y[p] = y;
    Z[p] = z;
    if (ids) ids[p] = p + first_id;
}

// one particle: three sequential TSC displacements (src/mitos.c:827-863) then
// the periodic wrap (:1072-1077), in the reference's double arithmetic
template <typename T>
__device__ __forceinline__ void displace_one(const T* __restrict__ px, const T* __restrict__ py,
                                             const T* __restrict__ pz, unsigned rs, int N,
                                             double boxlen, float& x, float& y, float& z) {
#pragma clang fp contract(off)
    const float invN = 1.0f / N;
    double d = tsc_gather<T, double, false>(px, rs, N, invN, boxlen, x, y, z);
    x = (float)((double)x - d);
    d = tsc_gather<T, double, false>(py, rs, N, invN, boxlen, x, y, z);
    y = (float)((double)y - d);
    d = tsc_gather<T, double, false>(pz, rs, N, invN, boxlen, x, y, z);
    z = (float)((double)z - d);
    x = (float)fwrap_fast(x, boxlen);
    y = (float)fwrap_fast(y, boxlen);
    z = (float)fwrap_fast(z, boxlen);
}

// particles at X/Y/Z displaced through grids of row stride rs (no halo)
template <typename T>
__global__ void __launch_bounds__(256)
    k_displace(const T* __restrict__ px, const T* __restrict__ py, const T* __restrict__ pz,
               unsigned rs, int N, double boxlen, long long npart, float* X, float* Y, float* Z) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npart) return;
    float x = X[p], y = Y[p], z = Z[p];
    displace_one<T>(px, py, pz, rs, N, boxlen, x, y, z);
    X[p] = x;
    Y[p] = y;
    Z[p] = z;
}

// ---- displacement grids with a periodic z halo -------------------------------------
// The pipeline stores each displacement grid as rows of rs >= N + MITOS_PSI_HO + 1
// values: value (x, y, z) at row (x*N + y) offset MITOS_PSI_HO + z, with the
// halo values z = -1 and z = N (copies of z = N-1 and z = 0) beside the row, so
// the three z taps of a TSC stencil are always three consecutive values.
#define MITOS_PSI_HO 4
// The stencil form (the lattice moves read only values at a lattice z, never
// a z tap) has no halo: value z at offset MITOS_PSI_SO + z = z, so the 64
// values of a wave's z chunk are two whole 128-byte lines, not three.
#define MITOS_PSI_SO 0

// FP32 TSC weights: nearest cell c = floor(X + 1/2), u = X - c in [-1/2, 1/2],
// weights of cells c-1, c, c+1.  Same cells and (up to rounding) the same
// weights as tsc_axis.
__device__ __forceinline__ void tsc_axis_near(float X, int& c, float w[3]) {
    const float cf = floorf(X + 0.5f);
    const float u = X - cf;
    c = (int)cf;
    const float a = 0.5f - u, b = 0.5f + u;
    w[0] = 0.5f * (a * a);
    w[1] = 0.75f - u * u;
    w[2] = 0.5f * (b * b);
}

// Byte offsets of the x and y taps c-1, c, c+1 of a wrapped centre cell
__device__ __forceinline__ void tap_rows(int cb, int N, unsigned stride, unsigned o[3]) {
    const int m = cb == 0 ? N - 1 : cb - 1;
    const int p = cb == N - 1 ? 0 : cb + 1;
    o[0] = (unsigned)m * stride;
    o[1] = (unsigned)cb * stride;
    o[2] = (unsigned)p * stride;
}

// TSC interpolation from a halo-padded grid.  REF = true: the weights and the
// 27-term sum of gridTSC_dg (src/grids_interp.c:267-309) with every operation
// rounded as there.  REF = false (float): separable weights and fused sums.
// Nearest cell of a position in cell units given in double, and the three
// TSC weights in M from the offset u = X - c (exact in double, then rounded
// once), so a float gather weights its taps at the position the reference
// forms in double (x*N/boxlen, src/grids_interp.c:269-271).
template <typename M>
__device__ __forceinline__ void tsc_axis_near_pos(double X, int& c, M w[3]) {
    const double cf = floor(X + 0.5);
    const M u = (M)(X - cf);
    c = (int)cf;
    const M a = (M)0.5 - u, b = (M)0.5 + u;
    w[0] = (M)0.5 * (a * a);
    w[1] = (M)0.75 - u * u;
    w[2] = (M)0.5 * (b * b);
}

template <typename T, typename M, bool REF>
__device__ __forceinline__ M tsc_halo(const T* __restrict__ box, unsigned rs, int N, float invN,
                                      double X, double Y, double Z) {
#pragma clang fp contract(off)
    int cx, cy, cz;
    M wx[3], wy[3], wz[3];
    if constexpr (REF) {
        tsc_axis(X, cx, wx);
        tsc_axis(Y, cy, wy);
        tsc_axis(Z, cz, wz);
        cx++;
        cy++;
        cz++;
    } else {
        tsc_axis_near_pos(X, cx, wx);
        tsc_axis_near_pos(Y, cy, wy);
        tsc_axis_near_pos(Z, cz, wz);
    }
    const int xb = wrapf(cx, N, invN), yb = wrapf(cy, N, invN), zb = wrapf(cz, N, invN);
    unsigned ox[3], oy[3];
    tap_rows(xb, N, (unsigned)N * rs, ox);
    tap_rows(yb, N, rs, oy);
    const T* zp = box + (MITOS_PSI_HO - 1) + zb;  // value z = zb - 1 of row 0
    M sum = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        M sa = 0;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const T* r = zp + (ox[a] + oy[b]);
            const M v0 = (M)r[0], v1 = (M)r[1], v2 = (M)r[2];
            if constexpr (REF) {
                const M wab = wx[a] * wy[b];
                sum += v0 * (wab * wz[0]);
                sum += v1 * (wab * wz[1]);
                sum += v2 * (wab * wz[2]);
            } else {
                const M sr = fmaf(wz[2], v2, fmaf(wz[1], v1, wz[0] * v0));
                sa = fmaf(wy[b], sr, sa);
            }
        }
        if constexpr (!REF) sum = fmaf(wx[a], sa, sum);
    }
    return sum;
}

// src/mitos.c:827-863 then :1072-1077 for the particles of an Ml^3 lattice
// (src/particle.c:87-129, spacing (float)L / Ml) over halo-padded
// displacement grids of N^3 cells.  One work-group per (lattice row, chunk of
// NT sites along z); the FP64 instance (M = double) keeps the reference's
// arithmetic, the FP32 one (M = float) moves particles in float.
template <typename T, typename M, int NT>
__global__ void __launch_bounds__(NT)
    k_displace_lattice(const T* __restrict__ px, const T* __restrict__ py,
                       const T* __restrict__ pz, unsigned rs, int N, int Ml, int nch,
                       double boxlen, float spacing, float* __restrict__ X,
                       float* __restrict__ Y,
                       float* __restrict__ Z) {
#pragma clang fp contract(off)
    constexpr bool REF = sizeof(M) == 8;
    const unsigned bid = xcd_block(blockIdx.x, gridDim.x);
    const unsigned row = bid / (unsigned)nch;
    const int z = (int)(bid - row * (unsigned)nch) * NT + threadIdx.x;
    if (z >= Ml) return;
    const int qx = (int)(row / (unsigned)Ml), qy = (int)(row - (unsigned)qx * Ml);
    float x = qx * spacing, y = qy * spacing, zz = z * spacing;  // src/particle.c:107-109
    const float invN = 1.0f / N;
    if constexpr (REF) {
        const double bl = boxlen;
        double d = tsc_halo<T, double, true>(px, rs, N, invN, (double)x * N / bl,
                                             (double)y * N / bl, (double)zz * N / bl);
        x = (float)((double)x - d);
        d = tsc_halo<T, double, true>(py, rs, N, invN, (double)x * N / bl, (double)y * N / bl,
                                      (double)zz * N / bl);
        y = (float)((double)y - d);
        d = tsc_halo<T, double, true>(pz, rs, N, invN, (double)x * N / bl, (double)y * N / bl,
                                      (double)zz * N / bl);
        zz = (float)((double)zz - d);
    } else {
        // cell coordinates as the reference forms them (double x*N/boxlen);
        // weights, sums and moves in float
        const double bl = boxlen;
        float d = tsc_halo<T, float, false>(px, rs, N, invN, (double)x * N / bl,
                                            (double)y * N / bl, (double)zz * N / bl);
        x = x - d;
        d = tsc_halo<T, float, false>(py, rs, N, invN, (double)x * N / bl, (double)y * N / bl,
                                      (double)zz * N / bl);
        y = y - d;
        d = tsc_halo<T, float, false>(pz, rs, N, invN, (double)x * N / bl, (double)y * N / bl,
                                      (double)zz * N / bl);
        zz = zz - d;
    }
    const long long p = (long long)row * Ml + z;
    X[p] = (float)fwrap_fast(x, boxlen);
    Y[p] = (float)fwrap_fast(y, boxlen);
    Z[p] = (float)fwrap_fast(zz, boxlen);
}

// FP32 lattice displacement.  For a coordinate on the lattice the TSC window
// of src/grids_interp.c:291-296 has the three weights (1/8, 3/4, 1/8), so the
// sequential moves of src/mitos.c:827-863 reduce to
//   x -= (S_xyz psi_x)(q),  y -= TSC_x(S_yz psi_y)(x', q_y, q_z),
//   z -= TSC_xy(S_z psi_z)(x', y', q_z),
// with S_a the 3-tap lattice stencil along axis a.  The stencils are applied
// in k-space when the displacement grids are made (factor 3/4 + cos(2 pi n/N)/4
// per axis), so the moves read 1 + 3 + 9 values instead of 3 x 27.
// Work-groups walk strips of TX x-planes row by row in y, so the x taps of
// neighbouring rows stay in L2.
// TSC weights in arithmetic M (float: tsc_axis_near; double: the same
// closed form in double) and, for the adjoint, their derivatives in X.
__device__ __forceinline__ void tsc_axis_near(double X, int& c, double w[3]) {
    const double cf = floor(X + 0.5);
    const double u = X - cf;
    c = (int)cf;
    const double a = 0.5 - u, b = 0.5 + u;
    w[0] = 0.5 * (a * a);
    w[1] = 0.75 - u * u;
    w[2] = 0.5 * (b * b);
}
template <typename M>
__device__ __forceinline__ void tsc_axis_near_d(M X, int& c, M w[3], M dw[3]) {
    const M cf = floor(X + (M)0.5);
    const M u = X - cf;
    c = (int)cf;
    const M a = (M)0.5 - u, b = (M)0.5 + u;
    w[0] = (M)0.5 * (a * a);
    w[1] = (M)0.75 - u * u;
    w[2] = (M)0.5 * (b * b);
    dw[0] = -a;
    dw[1] = (M)-2 * u;
    dw[2] = b;
}

// T: grid and arithmetic type (float in the FP32 pipeline; double when the
// FP64 pipeline runs in this form for the adjoint).  Each thread moves PPT
// particles of consecutive lattice rows at the same z, phase by phase, so
// the three dependent gathers of one particle overlap with the others'.
template <typename T, int NT, int TX, int PPT>
__global__ void __launch_bounds__(NT)
    k_displace_lattice_s(const T* __restrict__ sx, const T* __restrict__ sy,
                         const T* __restrict__ sz, unsigned rs, int N, int nch, T Lf,
                         float spacing, double boxlen, float* __restrict__ X,
                         float* __restrict__ Y, float* __restrict__ Z) {
    static_assert(TX == 1 || TX % PPT == 0, "a thread's rows lie in one strip");
    const unsigned bid = xcd_block(blockIdx.x, gridDim.x);
    const unsigned rowg = bid / (unsigned)nch;
    const int qz = (int)(bid - rowg * (unsigned)nch) * NT + threadIdx.x;
    if (qz >= N) return;
    const float invN = 1.0f / N;
    // cells per length in double: the moved coordinates' cells within a
    // double rounding of the reference's x*N/boxlen (float weights follow)
    const double cpl = (double)N / boxlen;
    const unsigned plane = (unsigned)N * rs;
    const unsigned zo = MITOS_PSI_SO + qz;
    int qx[PPT], qy[PPT];
    unsigned yz[PPT];
    float x[PPT], y[PPT], z[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        // TX = 1: PPT consecutive rows along y (the taps of a work-group then
        // stay in a few x-planes); TX > 1: strips of TX x-planes walked row
        // by row in y
        const unsigned rowi = rowg * PPT + j;
        const unsigned xi = TX > 1 ? rowi % TX : 0, r2 = TX > 1 ? rowi / TX : rowi;
        qy[j] = (int)(r2 % (unsigned)N);
        qx[j] = (int)(r2 / (unsigned)N) * TX + (int)xi;
        yz[j] = (unsigned)qy[j] * rs + zo;
        // src/particle.c:107-109
        x[j] = qx[j] * spacing;
        y[j] = qy[j] * spacing;
        z[j] = qz * spacing;
    }
    // x move: the stencilled x displacement at the site
    T sxv[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) sxv[j] = sx[(unsigned)qx[j] * plane + yz[j]];
#pragma unroll
    for (int j = 0; j < PPT; j++) x[j] = (float)((T)x[j] - sxv[j]);
    // y move: TSC along x at (x', q_y, q_z)
    T wx[PPT][3];
    unsigned ox[PPT][3];
    T syv[PPT][3];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        int cx;
        tsc_axis_near_pos<T>((double)x[j] * cpl, cx, wx[j]);
        tap_rows(wrapf(cx, N, invN), N, plane, ox[j]);
#pragma unroll
        for (int a = 0; a < 3; a++) syv[j][a] = sy[ox[j][a] + yz[j]];
    }
#pragma unroll
    for (int j = 0; j < PPT; j++)
        y[j] = (float)((T)y[j] -
                       fma(wx[j][2], syv[j][2], fma(wx[j][1], syv[j][1], wx[j][0] * syv[j][0])));
    // z move: TSC along x and y at (x', y', q_z)
    T wy[PPT][3];
    T szv[PPT][9];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        int cy;
        tsc_axis_near_pos<T>((double)y[j] * cpl, cy, wy[j]);
        unsigned oy[3];
        tap_rows(wrapf(cy, N, invN), N, rs, oy);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) szv[j][a * 3 + b] = sz[ox[j][a] + oy[b] + zo];
    }
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        T d = 0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T r = fma(wy[j][2], szv[j][a * 3 + 2],
                            fma(wy[j][1], szv[j][a * 3 + 1], wy[j][0] * szv[j][a * 3]));
            d = fma(wx[j][a], r, d);
        }
        z[j] = (float)((T)z[j] - d);
        const long long p = ((long long)qx[j] * N + qy[j]) * N + qz;
        const float Lp = (float)Lf;
        X[p] = fwrap_float(x[j], Lp, boxlen);
        Y[p] = fwrap_float(y[j], Lp, boxlen);
        Z[p] = fwrap_float(z[j], Lp, boxlen);
    }
}

// wave-wide minimum and maximum of an int
__device__ __forceinline__ void wave_minmax(int& lo, int& hi) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const int a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
}

// k_displace_lattice_s with the taps staged in LDS.  One wave per work-group
// moves PPT x-consecutive lattice rows (one y) over 64 consecutive z.  After
// the x move the wave's x taps span a few rows [lo - 1, hi + 1]: those rows of
// the y displacement grid (the 64 z of the wave) are loaded whole, as
// coalesced 256-byte rows, into LDS, and the y move reads its taps there;
// the rows (x taps) x (y taps) of the z displacement grid follow after the y
// move.  A wave whose taps span more than MAXX x rows or MAXY y rows reads
// them from the grids directly.  Weights, taps and the order of the sums are
// those of k_displace_lattice_s, so the positions are the same bits.
// N must be a multiple of 64; `staged` = false reads every tap directly.
template <typename T, int TX, int PPT, int MAXX, int MAXY>
__global__ void __launch_bounds__(64)
    k_displace_lattice_w(const T* __restrict__ sx, const T* __restrict__ sy,
                         const T* __restrict__ sz, unsigned rs, int N, T Lf, float spacing,
                         double boxlen, float* __restrict__ X, float* __restrict__ Y,
                         float* __restrict__ Z, int staged) {
    static_assert(TX % PPT == 0, "a thread's rows lie in one strip");
    __shared__ T buf[MAXX * MAXY][64];
    const int nch = N / 64;
    const unsigned bid = xcd_block(blockIdx.x, gridDim.x);
    const unsigned rowg = bid / (unsigned)nch;
    const int lane = threadIdx.x;
    const int qz = (int)(bid - rowg * (unsigned)nch) * 64 + lane;
    const float invN = 1.0f / N;
    const double cpl = (double)N / boxlen;  // as in k_displace_lattice_s
    const unsigned plane = (unsigned)N * rs;
    const unsigned zo = MITOS_PSI_SO + qz;
    int qx[PPT];
    float x[PPT], y[PPT], z[PPT];
    const unsigned row0 = rowg * PPT;
    const unsigned r2 = row0 / TX;
    const int qy = (int)(r2 % (unsigned)N);
    const unsigned yz = (unsigned)qy * rs + zo;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        qx[j] = (int)(r2 / (unsigned)N) * TX + (int)((row0 + j) % TX);
        // src/particle.c:107-109
        x[j] = qx[j] * spacing;
        y[j] = qy * spacing;
        z[j] = qz * spacing;
    }
    // x move: the stencilled x displacement at the site
    T sxv[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) sxv[j] = sx[(unsigned)qx[j] * plane + yz];
#pragma unroll
    for (int j = 0; j < PPT; j++) x[j] = (float)((T)x[j] - sxv[j]);
    // y move: TSC along x at (x', q_y, q_z)
    T wx[PPT][3];
    int cx[PPT];
    int lo = 1 << 30, hi = -(1 << 30);
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        tsc_axis_near_pos<T>((double)x[j] * cpl, cx[j], wx[j]);
        lo = cx[j] < lo ? cx[j] : lo;
        hi = cx[j] > hi ? cx[j] : hi;
    }
    wave_minmax(lo, hi);
    const int nx = hi - lo + 3;  // x rows lo - 1 .. hi + 1
    const bool stage_x = staged && nx <= MAXX;
    T syv[PPT][3];
    if (stage_x) {
        T v[MAXX];
#pragma unroll
        for (int i = 0; i < MAXX; i++)
            if (i < nx) v[i] = sy[(unsigned)wrapf(lo - 1 + i, N, invN) * plane + yz];
#pragma unroll
        for (int i = 0; i < MAXX; i++)
            if (i < nx) buf[i][lane] = v[i];
#pragma unroll
        for (int j = 0; j < PPT; j++)
#pragma unroll
            for (int a = 0; a < 3; a++) syv[j][a] = buf[cx[j] - lo + a][lane];
    } else {
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            unsigned o[3];
            tap_rows(wrapf(cx[j], N, invN), N, plane, o);
#pragma unroll
            for (int a = 0; a < 3; a++) syv[j][a] = sy[o[a] + yz];
        }
    }
#pragma unroll
    for (int j = 0; j < PPT; j++)
        y[j] = (float)((T)y[j] -
                       fma(wx[j][2], syv[j][2], fma(wx[j][1], syv[j][1], wx[j][0] * syv[j][0])));
    // z move: TSC along x and y at (x', y', q_z)
    T wy[PPT][3];
    int cy[PPT];
    int ylo = 1 << 30, yhi = -(1 << 30);
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        tsc_axis_near_pos<T>((double)y[j] * cpl, cy[j], wy[j]);
        ylo = cy[j] < ylo ? cy[j] : ylo;
        yhi = cy[j] > yhi ? cy[j] : yhi;
    }
    wave_minmax(ylo, yhi);
    const int ny = yhi - ylo + 3;
    T szv[PPT][9];
    if (stage_x && ny <= MAXY) {
        // rows (x lo - 1 + i, y ylo - 1 + k) at buf[i * MAXY + k]; the y
        // move's reads of buf precede these writes in the wave's order
        unsigned oyk[MAXY];
#pragma unroll
        for (int k = 0; k < MAXY; k++) oyk[k] = (unsigned)wrapf(ylo - 1 + k, N, invN) * rs + zo;
        for (int i = 0; i < nx; i += 2) {
            const unsigned ox0 = (unsigned)wrapf(lo - 1 + i, N, invN) * plane;
            const unsigned ox1 = (unsigned)wrapf(lo + i, N, invN) * plane;
            T v[2 * MAXY];
#pragma unroll
            for (int k = 0; k < MAXY; k++) {
                if (k < ny) v[k] = sz[ox0 + oyk[k]];
                if (k < ny && i + 1 < nx) v[MAXY + k] = sz[ox1 + oyk[k]];
            }
#pragma unroll
            for (int k = 0; k < MAXY; k++) {
                if (k < ny) buf[i * MAXY + k][lane] = v[k];
                if (k < ny && i + 1 < nx) buf[(i + 1) * MAXY + k][lane] = v[MAXY + k];
            }
        }
#pragma unroll
        for (int j = 0; j < PPT; j++)
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    szv[j][a * 3 + b] = buf[(cx[j] - lo + a) * MAXY + (cy[j] - ylo + b)][lane];
    } else {
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            unsigned ox[3], oy[3];
            tap_rows(wrapf(cx[j], N, invN), N, plane, ox);
            tap_rows(wrapf(cy[j], N, invN), N, rs, oy);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) szv[j][a * 3 + b] = sz[ox[a] + oy[b] + zo];
        }
    }
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        T d = 0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T r = fma(wy[j][2], szv[j][a * 3 + 2],
                            fma(wy[j][1], szv[j][a * 3 + 1], wy[j][0] * szv[j][a * 3]));
            d = fma(wx[j][a], r, d);
        }
        z[j] = (float)((T)z[j] - d);
        const long long p = ((long long)qx[j] * N + qy) * N + qz;
        const float Lp = (float)Lf;
        X[p] = fwrap_float(x[j], Lp, boxlen);
        Y[p] = fwrap_float(y[j], Lp, boxlen);
        Z[p] = fwrap_float(z[j], Lp, boxlen);
    }
}

// work-group minimum and maximum of an int over NW waves (`red` holds the
// waves' pairs)
template <int NW>
__device__ __forceinline__ void block_minmax(int& lo, int& hi, int (*red)[2]) {
    wave_minmax(lo, hi);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = lo;
        red[wv][1] = hi;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < NW; w++) {
        lo = red[w][0] < lo ? red[w][0] : lo;
        hi = red[w][1] > hi ? red[w][1] : hi;
    }
}

// FP32 lattice displacement over TXN x TY x TZ lattice tiles (x, y, z), one
// work-group of (TXN / PP) * TY * TZ threads per tile; thread t moves the PP
// x-consecutive particles (x0 + PP (t / (TY TZ)) + j, y0 + (t / TZ) % TY,
// z0 + t % TZ).  The taps of a tile's y and z moves lie in small windows of
// the displacement grids (the displacement varies little over a compact
// tile): after the x move the tile's x taps span [xlo - 1, xhi + 1].  Those
// rows of S_yz psi_y (the tile's TY y and TZ z) are loaded by the whole
// work-group into LDS, at most U loads in flight per thread, and the y move
// reads its taps there; after the y move the rows (x window) x (y window) of
// S_z psi_z follow the same way.  A window over MAXX x rows or MAXR rows is
// read from the grids directly.  The weights, taps and order of the sums are
// those of k_displace_lattice_s, so the positions are the same bits.
// ORDER 0: tiles walk z, then x, then y; 1: z, y, x; 2: z, then y within
// blocks of 4, then x, then the y blocks.  N must be a multiple of TZ.
template <typename T, int TXN, int TY, int TZ, int PP, int MAXX, int MAXR, int U, int ORDER>
__global__ void __launch_bounds__(TXN / PP * TY * TZ)
    k_displace_tile(const T* __restrict__ sx, const T* __restrict__ sy,
                    const T* __restrict__ sz, unsigned rs, int N, T Lf, float spacing,
                    double boxlen, float* __restrict__ X, float* __restrict__ Y,
                    float* __restrict__ Z, int staged) {
    static_assert(MAXX * TY <= MAXR, "the y window fits the buffer");
    constexpr int NT = TXN / PP * TY * TZ, NW = NT / 64;
    constexpr int XG = NT / (TY * TZ);  // x groups of PP rows
    constexpr int RS = NT / TZ;        // window rows per round of the work-group
    __shared__ T buf[(MAXR + 1) * TZ];  // window rows and a spare row
    __shared__ unsigned roff[MAXR];      // grid offsets of the z move's window rows
    __shared__ int redx[NW][2], redy[NW][2];
    const int t = threadIdx.x;
    const int zl = t % TZ, yl = (t / TZ) % TY;
    // x group of the thread: a whole number of waves (TY * TZ a multiple of
    // 64), so wave-uniform and kept in a scalar register
    static_assert((TY * TZ) % 64 == 0, "an x group is whole waves");
    const int xh = __builtin_amdgcn_readfirstlane(t / (TY * TZ));
    const unsigned nzt = (unsigned)N / TZ, nxt = (unsigned)N / TXN, nyt = (unsigned)N / TY;
    const unsigned bid = xcd_block(blockIdx.x, gridDim.x);
    const unsigned zt = bid % nzt, r1 = bid / nzt;
    unsigned xt, yt;
    if constexpr (ORDER == 1) {
        xt = r1 / nyt;
        yt = r1 % nyt;
    } else if constexpr (ORDER == 2) {
        const unsigned r2 = r1 / 4u;
        xt = r2 % nxt;
        yt = (r2 / nxt) * 4u + r1 % 4u;
    } else {
        xt = r1 % nxt;
        yt = r1 / nxt;
    }
    const int qz = (int)zt * TZ + zl, qy = (int)yt * TY + yl;
    const int x0 = (int)xt * TXN + xh * PP;
    const float invN = 1.0f / N;
    const double cpl = (double)N / boxlen;  // as in k_displace_lattice_s
    const unsigned plane = (unsigned)N * rs;
    float x[PP], y[PP];
    const float z = qz * spacing;  // src/particle.c:107-109
    T sxv[PP];
#pragma unroll
    for (int j = 0; j < PP; j++) {
        x[j] = (x0 + j) * spacing;
        y[j] = qy * spacing;
        sxv[j] = sx[(unsigned)(x0 + j) * plane + (unsigned)qy * rs + qz];
    }
    // x move: the stencilled x displacement at the site
#pragma unroll
    for (int j = 0; j < PP; j++) x[j] = (float)((T)x[j] - sxv[j]);
    T wx[PP][3];
    int cx[PP];
    int lo = 1 << 30, hi = -(1 << 30);
#pragma unroll
    for (int j = 0; j < PP; j++) {
        tsc_axis_near_pos<T>((double)x[j] * cpl, cx[j], wx[j]);
        lo = cx[j] < lo ? cx[j] : lo;
        hi = cx[j] > hi ? cx[j] : hi;
    }
    block_minmax<NW>(lo, hi, redx);
    // work-group uniform: kept in scalar registers
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);
    const int nx = hi - lo + 3;  // x rows lo - 1 .. hi + 1
    // (windows no wider than the grid keep the incremental wraps below to
    // one correction each way)
    const bool stage_x = staged && nx <= MAXX && nx <= N;
    // y move: TSC along x at (x', q_y, q_z)
    T syv[PP][3];
    if (stage_x) {
        // window row xi (x = lo - 1 + xi) of this thread's (y, z) column at
        // buf[(xi * TY + yl) * TZ + zl]; the thread loads rows xh, xh + XG,
        // ...; its wrapped x follows incrementally
        const unsigned col = (unsigned)qy * rs + qz;
        int xw = wrapf(lo - 1 + xh, N, invN);
        for (int k0 = xh; k0 < nx; k0 += XG * U) {
            T v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                // rows past the window are not loaded (a wave-uniform test)
                if (k0 + XG * u < nx) v[u] = sy[__umul24((unsigned)xw, plane) + col];
                xw += XG;
                if (xw >= N) xw -= N;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int xi = k0 + XG * u;
                if (xi < nx) buf[(xi * TY + yl) * TZ + zl] = v[u];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PP; j++)
#pragma unroll
            for (int a = 0; a < 3; a++) syv[j][a] = buf[((cx[j] - lo + a) * TY + yl) * TZ + zl];
    } else {
#pragma unroll
        for (int j = 0; j < PP; j++) {
            unsigned o[3];
            tap_rows(wrapf(cx[j], N, invN), N, plane, o);
#pragma unroll
            for (int a = 0; a < 3; a++) syv[j][a] = sy[o[a] + (unsigned)qy * rs + qz];
        }
    }
#pragma unroll
    for (int j = 0; j < PP; j++)
        y[j] = (float)((T)y[j] -
                       fma(wx[j][2], syv[j][2], fma(wx[j][1], syv[j][1], wx[j][0] * syv[j][0])));
    // z move: TSC along x and y at (x', y', q_z)
    T wy[PP][3];
    int cy[PP];
    int ylo = 1 << 30, yhi = -(1 << 30);
#pragma unroll
    for (int j = 0; j < PP; j++) {
        tsc_axis_near_pos<T>((double)y[j] * cpl, cy[j], wy[j]);
        ylo = cy[j] < ylo ? cy[j] : ylo;
        yhi = cy[j] > yhi ? cy[j] : yhi;
    }
    // (the barrier inside also ends the y move's reads of buf)
    block_minmax<NW>(ylo, yhi, redy);
    ylo = __builtin_amdgcn_readfirstlane(ylo);
    yhi = __builtin_amdgcn_readfirstlane(yhi);
    const int ny = yhi - ylo + 3;
    T szv[PP][9];
    if (stage_x && nx * ny <= MAXR && ny <= N) {
        // window row r = xi * ny + yi (x = lo - 1 + xi, y = ylo - 1 + yi) at
        // buf[r * TZ + zl]; the work-group first tabulates every row's
        // wrapped grid offset in LDS, then the TZ-lane group g = t / TZ loads
        // the g-th of RS blocks of consecutive rows (a load past the block
        // rereads its last row and is stored to the spare row MAXR)
        const int nrow = nx * ny;
        const float rny = 1.0f / (float)ny;
        for (int r = t; r < nrow; r += NT) {
            int xi = (int)((float)r * rny);
            int yi = r - xi * ny;
            if (yi < 0) {
                xi--;
                yi += ny;
            } else if (yi >= ny) {
                xi++;
                yi -= ny;
            }
            roff[r] = __umul24((unsigned)wrapf(lo - 1 + xi, N, invN), plane) +
                      __umul24((unsigned)wrapf(ylo - 1 + yi, N, invN), rs);
        }
        __syncthreads();
        const int per = (nrow + RS - 1) / RS;
        const int r0 = (t / TZ) * per;
        const int rend = r0 + per < nrow ? r0 + per : nrow;
        for (int r = r0; r < rend; r += U) {
            T v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int rr = r + u < rend ? r + u : rend - 1;
                v[u] = sz[roff[rr] + qz];
            }
#pragma unroll
            for (int u = 0; u < U; u++) buf[(r + u < rend ? r + u : MAXR) * TZ + zl] = v[u];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PP; j++)
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    szv[j][a * 3 + b] =
                        buf[((cx[j] - lo + a) * ny + (cy[j] - ylo + b)) * TZ + zl];
    } else {
#pragma unroll
        for (int j = 0; j < PP; j++) {
            unsigned ox[3], oy[3];
            tap_rows(wrapf(cx[j], N, invN), N, plane, ox);
            tap_rows(wrapf(cy[j], N, invN), N, rs, oy);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) szv[j][a * 3 + b] = sz[ox[a] + oy[b] + qz];
        }
    }
    const float Lp = (float)Lf;
#pragma unroll
    for (int j = 0; j < PP; j++) {
        T d = 0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const T rr = fma(wy[j][2], szv[j][a * 3 + 2],
                             fma(wy[j][1], szv[j][a * 3 + 1], wy[j][0] * szv[j][a * 3]));
            d = fma(wx[j][a], rr, d);
        }
        const float zz = (float)((T)z - d);
        const long long p = ((long long)(x0 + j) * N + qy) * N + qz;
        X[p] = fwrap_float(x[j], Lp, boxlen);
        Y[p] = fwrap_float(y[j], Lp, boxlen);
        Z[p] = fwrap_float(zz, Lp, boxlen);
    }
}

// FP64 lattice displacement in the reference's arithmetic (the 27-tap sums of
// tsc_halo<REF>, src/grids_interp.c:267-309, for the moves of
// src/mitos.c:827-863) over TXN x TY x TZ lattice tiles, one work-group of
// (TXN / PP) * TY * TZ threads per tile, thread t moving the PP x-consecutive
// particles (x0 + PP (t / (TY TZ)) + j, y0 + (t / TZ) % TY, z0 + t % TZ).
// Each move reads its 27 taps from a window of its halo-padded displacement
// grid staged in LDS by the whole work-group: window row r holds the ZW =
// TZ + 2 values z0 - 1 .. z0 + TZ of grid row (xs + r / nyr, ys + r % nyr)
// (wrapped), 32 lanes a row.  The x move's window is the tile's rows with
// one row of margin; the y move's spans the moved particles' x taps
// [xlo - 1, xhi + 1] times the tile's y rows with margin; the z move's spans
// those x taps times the moved y taps.  A window over MAXR rows is not
// staged (that move reads the grid), and a particle whose taps leave the
// window (not the case for lattice sites) reads the grid too; the weights,
// taps and order of the sums are those of tsc_halo<REF>, so the positions
// are the same bits as k_displace_lattice's.  N must be a multiple of TXN,
// TY and TZ.  Four waves per SIMD: the kernel fits 128 registers without
// spills (168 when the compiler is left free, three work-groups per CU).
template <int TXN, int TY, int TZ, int PP, int MAXR, int U>
__global__ void __launch_bounds__(TXN / PP * TY * TZ) __attribute__((amdgpu_waves_per_eu(4)))
    k_displace_tile_ref(const double* __restrict__ px, const double* __restrict__ py,
                        const double* __restrict__ pz, unsigned rs, int N, float spacing,
                        double boxlen, float* __restrict__ X, float* __restrict__ Y,
                        float* __restrict__ Z) {
#pragma clang fp contract(off)
    constexpr int NT = TXN / PP * TY * TZ, NW = NT / 64;
    constexpr int ZW = TZ + 2;  // z values of a window row
    constexpr int RG = NT / 32;  // rows staged per round, 32 lanes a row
    static_assert(ZW <= 32 && NT % 64 == 0, "a window row fits 32 lanes");
    static_assert((TXN + 2) * (TY + 2) <= MAXR, "the x move's window fits the buffer");
    __shared__ double buf[MAXR * ZW];
    __shared__ int redx[NW][2], redy[NW][2];
    const int t = threadIdx.x;
    const int zl = t % TZ, yl = (t / TZ) % TY, xh = t / (TY * TZ);
    const unsigned nzt = (unsigned)N / TZ, nxt = (unsigned)N / TXN;
    const unsigned bid = xcd_block(blockIdx.x, gridDim.x);
    const unsigned zt = bid % nzt, r1 = bid / nzt;
    const int z0 = (int)zt * TZ, y0 = (int)(r1 / nxt) * TY, xb = (int)(r1 % nxt) * TXN;
    const int qz = z0 + zl, qy = y0 + yl, x0 = xb + xh * PP;
    const float invN = 1.0f / N;
    const double bl = boxlen;
    const unsigned plane = (unsigned)N * rs;
    const int rl = t & 31, rg = t >> 5;
    // stage window rows 0 .. nrow - 1 of src, nyr rows per x, from (xs, ys)
    auto stage = [&](const double* __restrict__ src, int nrow, int nyr, int xs, int ys) {
        const float rcp = 1.0f / (float)nyr;
        for (int r0 = rg; r0 < nrow; r0 += RG * U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int r = r0 + RG * u;
                v[u] = 0;
                if (r < nrow && rl < ZW) {
                    int xi = (int)((float)r * rcp);
                    int yi = r - xi * nyr;
                    if (yi < 0) {
                        xi--;
                        yi += nyr;
                    } else if (yi >= nyr) {
                        xi++;
                        yi -= nyr;
                    }
                    const unsigned o = (unsigned)wrapf(xs + xi, N, invN) * plane +
                                       (unsigned)wrapf(ys + yi, N, invN) * rs +
                                       (unsigned)(MITOS_PSI_HO - 1 + z0 + rl);
                    v[u] = src[o];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int r = r0 + RG * u;
                if (r < nrow && rl < ZW) buf[r * ZW + rl] = v[u];
            }
        }
    };
    // the 27-term sum of tsc_halo<REF> with taps from the staged window (or
    // from the grid when they are not all in it)
    auto tap_sum = [&](const double* __restrict__ src, bool staged, int nxr, int nyr, int xs,
                       int ys, double Xc, double Yc, double Zc) -> double {
#pragma clang fp contract(off)
        int cx, cy, cz;
        double wx[3], wy[3], wz[3];
        tsc_axis(Xc, cx, wx);
        tsc_axis(Yc, cy, wy);
        tsc_axis(Zc, cz, wz);
        cx++;
        cy++;
        cz++;
        const int ax = cx - 1 - xs, ay = cy - 1 - ys, az = cz - z0;  // window index of tap 0
        const bool in = staged && ax >= 0 && ax + 2 < nxr && ay >= 0 && ay + 2 < nyr &&
                        az >= 0 && az + 2 < ZW;
        if (!in) return tsc_halo<double, double, true>(src, rs, N, invN, Xc, Yc, Zc);
        double sum = 0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const double* r = buf + ((ax + a) * nyr + (ay + b)) * ZW + az;
                const double wab = wx[a] * wy[b];
                sum += r[0] * (wab * wz[0]);
                sum += r[1] * (wab * wz[1]);
                sum += r[2] * (wab * wz[2]);
            }
        }
        return sum;
    };
    float x[PP], y[PP];
    const float z = qz * spacing;  // src/particle.c:107-109
    const double Zc = (double)z * N / bl;
#pragma unroll
    for (int j = 0; j < PP; j++) {
        x[j] = (x0 + j) * spacing;
        y[j] = qy * spacing;
    }
    // x move: the tile's rows with one row of margin
    {
        const int nyr = TY + 2, nrow = (TXN + 2) * nyr;
        stage(px, nrow, nyr, xb - 1, y0 - 1);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PP; j++) {
            const double d = tap_sum(px, true, TXN + 2, nyr, xb - 1, y0 - 1,
                                     (double)x[j] * N / bl, (double)y[j] * N / bl, Zc);
            x[j] = (float)((double)x[j] - d);
        }
    }
    // y move: the moved x taps times the tile's y rows with margin
    int lo = 1 << 30, hi = -(1 << 30);
#pragma unroll
    for (int j = 0; j < PP; j++) {
        int c;
        double w[3];
        tsc_axis((double)x[j] * N / bl, c, w);
        lo = c < lo ? c : lo;
        hi = c > hi ? c : hi;
    }
    // (the barrier inside also ends the x move's reads of buf)
    block_minmax<NW>(lo, hi, redx);
    lo = __builtin_amdgcn_readfirstlane(lo);  // first taps of the centres: x window lo .. hi + 2
    hi = __builtin_amdgcn_readfirstlane(hi);
    const int nx = hi - lo + 3;
    {
        const int nyr = TY + 2;
        const bool st = nx <= MAXR && nx * nyr <= MAXR;
        if (st) stage(py, nx * nyr, nyr, lo, y0 - 1);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PP; j++) {
            const double d = tap_sum(py, st, nx, nyr, lo, y0 - 1, (double)x[j] * N / bl,
                                     (double)y[j] * N / bl, Zc);
            y[j] = (float)((double)y[j] - d);
        }
    }
    // z move: the x taps times the moved y taps
    int ylo = 1 << 30, yhi = -(1 << 30);
#pragma unroll
    for (int j = 0; j < PP; j++) {
        int c;
        double w[3];
        tsc_axis((double)y[j] * N / bl, c, w);
        ylo = c < ylo ? c : ylo;
        yhi = c > yhi ? c : yhi;
    }
    block_minmax<NW>(ylo, yhi, redy);
    ylo = __builtin_amdgcn_readfirstlane(ylo);
    yhi = __builtin_amdgcn_readfirstlane(yhi);
    const int ny = yhi - ylo + 3;
    float zz[PP];
    {
        const bool st = nx <= MAXR && ny <= MAXR && nx * ny <= MAXR;
        if (st) stage(pz, nx * ny, ny, lo, ylo);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PP; j++) {
            const double d = tap_sum(pz, st, nx, ny, lo, ylo, (double)x[j] * N / bl,
                                     (double)y[j] * N / bl, Zc);
            zz[j] = (float)((double)z - d);
        }
    }
#pragma unroll
    for (int j = 0; j < PP; j++) {
        const long long p = ((long long)(x0 + j) * N + qy) * N + qz;
        X[p] = (float)fwrap_fast(x[j], boxlen);
        Y[p] = (float)fwrap_fast(y[j], boxlen);
        Z[p] = (float)fwrap_fast(zz[j], boxlen);
    }
}

// analyse/src/mitos_read.c:285-320: base cell and the eight CIC weights of a
// particle, the weights as 2^-40 fixed point in the reference's corner order
// (000, 100, 010, 001, 110, 101, 011, 111).
__device__ __forceinline__ void cic_cell(float X, float Y, float Z, int N, double boxlen, int& iX,
                                         int& iY, int& iZ, double& dx, double& dy, double& dz) {
#pragma clang fp contract(off)
    const double cell = boxlen / N;
    const double Xc = fwrap_fast((double)X, boxlen) / cell;
    const double Yc = fwrap_fast((double)Y, boxlen) / cell;
    const double Zc = fwrap_fast((double)Z, boxlen) / cell;
    iX = (int)floor(Xc);
    iY = (int)floor(Yc);
    iZ = (int)floor(Zc);
    dx = Xc - iX;
    dy = Yc - iY;
    dz = Zc - iZ;
}

__device__ __forceinline__ void cic_weights(double dx, double dy, double dz,
                                            unsigned long long w[8]) {
#pragma clang fp contract(off)
    const double tx = 1.0 - dx, ty = 1.0 - dy, tz = 1.0 - dz;
    const double S = (double)(1ULL << MITOS_CIC_SHIFT);
    w[0] = (unsigned long long)(tx * ty * tz * S + 0.5);
    w[1] = (unsigned long long)(dx * ty * tz * S + 0.5);
    w[2] = (unsigned long long)(tx * dy * tz * S + 0.5);
    w[3] = (unsigned long long)(tx * ty * dz * S + 0.5);
    w[4] = (unsigned long long)(dx * dy * tz * S + 0.5);
    w[5] = (unsigned long long)(dx * ty * dz * S + 0.5);
    w[6] = (unsigned long long)(tx * dy * dz * S + 0.5);
    w[7] = (unsigned long long)(dx * dy * dz * S + 0.5);
}

// Float variant of cic_cell/cic_weights (FP32 pipeline): the offset in the
// cell is formed as X - iX * cell with one rounding, so it keeps the float
// position's accuracy; weights in units of 2^-24 (the float mantissa), so a
// 32-bit cell holds the mass of 256 particles.
__device__ __forceinline__ void cic_cell_f(float X, float Y, float Z, int N, float cell,
                                           float cinv, double boxlen, int& iX, int& iY, int& iZ,
                                           float& dx, float& dy, float& dz) {
    const float Lf = (float)boxlen;
    const float P[3] = {fwrap_float(X, Lf, boxlen), fwrap_float(Y, Lf, boxlen),
                        fwrap_float(Z, Lf, boxlen)};
    int I[3];
    float D[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int i = (int)floorf(P[a] * cinv);
        float r = fmaf(-(float)i, cell, P[a]);
        if (r < 0.0f) {
            i--;
            r += cell;
        } else if (r >= cell) {
            i++;
            r -= cell;
        }
        I[a] = i;
        D[a] = fminf(r * cinv, 1.0f);
    }
    iX = I[0];
    iY = I[1];
    iZ = I[2];
    dx = D[0];
    dy = D[1];
    dz = D[2];
}

__device__ __forceinline__ void cic_weights(float dx, float dy, float dz,
                                            unsigned long long w[8]) {
    const float tx = 1.0f - dx, ty = 1.0f - dy, tz = 1.0f - dz;
    const float S = (float)(1u << MITOS_CIC_SHIFT_F32);
    const float a = tx * ty, b = dx * ty, c = tx * dy, d = dx * dy;
    w[0] = (unsigned)fmaf(a * tz, S, 0.5f);
    w[1] = (unsigned)fmaf(b * tz, S, 0.5f);
    w[2] = (unsigned)fmaf(c * tz, S, 0.5f);
    w[3] = (unsigned)fmaf(a * dz, S, 0.5f);
    w[4] = (unsigned)fmaf(d * tz, S, 0.5f);
    w[5] = (unsigned)fmaf(b * dz, S, 0.5f);
    w[6] = (unsigned)fmaf(c * dz, S, 0.5f);
    w[7] = (unsigned)fmaf(d * dz, S, 0.5f);
}

// eight global atomic adds at wrapped base cell (iX, iY, iZ)
__device__ __forceinline__ void cic_global(unsigned long long* grid, int N, int iX, int iY, int iZ,
                                           const unsigned long long w[8]) {
    const long x0 = wrapi(iX, N), x1 = wrap1((int)x0 + 1, N);
    const long y0 = wrapi(iY, N), y1 = wrap1((int)y0 + 1, N);
    const long z0 = wrapi(iZ, N), z1 = wrap1((int)z0 + 1, N);
    atomicAdd(grid + (x0 * N + y0) * N + z0, w[0]);
    atomicAdd(grid + (x1 * N + y0) * N + z0, w[1]);
    atomicAdd(grid + (x0 * N + y1) * N + z0, w[2]);
    atomicAdd(grid + (x0 * N + y0) * N + z1, w[3]);
    atomicAdd(grid + (x1 * N + y1) * N + z0, w[4]);
    atomicAdd(grid + (x1 * N + y0) * N + z1, w[5]);
    atomicAdd(grid + (x0 * N + y1) * N + z1, w[6]);
    atomicAdd(grid + (x1 * N + y1) * N + z1, w[7]);
}

// analyse/src/mitos_read.c:281-334 one thread per particle, global atomics
// (any particle order; used when the lattice tiling below does not apply)
// (W as in k_cic_lattice)
template <typename W = double>
__global__ void __launch_bounds__(256)
    k_cic(const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ Z,
          long long npart, int N, double boxlen, unsigned long long* grid) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= npart) return;
    int iX, iY, iZ;
    W dx, dy, dz;
    if constexpr (sizeof(W) == 8)
        cic_cell(X[l], Y[l], Z[l], N, boxlen, iX, iY, iZ, dx, dy, dz);
    else
        cic_cell_f(X[l], Y[l], Z[l], N, (float)(boxlen / N), (float)(N / boxlen), boxlen, iX, iY,
                   iZ, dx, dy, dz);
    unsigned long long w[8];
    cic_weights(dx, dy, dz, w);
    cic_global(grid, N, iX, iY, iZ, w);
}

// The three stencil-form moves of one lattice particle in arithmetic T,
// without rounding the positions to float (used by the adjoint and by the
// forward run it differentiates).  Outputs the moved, unwrapped position.
template <typename T>
__device__ __forceinline__ void moves_s(const T* __restrict__ sx, const T* __restrict__ sy,
                                        const T* __restrict__ sz, unsigned rs, int N, T sN,
                                        float spacing, int qx, int qy, int qz, T& x, T& y,
                                        T& z) {
    const float invN = 1.0f / N;
    const unsigned plane = (unsigned)N * rs;
    const unsigned yz = (unsigned)qy * rs + MITOS_PSI_SO + qz;
    x = (T)(qx * spacing) - sx[(unsigned)qx * plane + yz];
    y = (T)(qy * spacing);
    z = (T)(qz * spacing);
    int cx;
    T wx[3];
    tsc_axis_near(x * sN, cx, wx);
    unsigned ox[3];
    tap_rows(wrapf(cx, N, invN), N, plane, ox);
    y -= wx[0] * sy[ox[0] + yz] + wx[1] * sy[ox[1] + yz] + wx[2] * sy[ox[2] + yz];
    int cy;
    T wy[3];
    tsc_axis_near(y * sN, cy, wy);
    unsigned oy[3];
    tap_rows(wrapf(cy, N, invN), N, rs, oy);
    const unsigned zo = MITOS_PSI_SO + qz;
    T d = 0;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) d += wx[a] * wy[b] * sz[ox[a] + oy[b] + zo];
    z -= d;
}

// periodic wrap of a coordinate in T (exact subtractions for |x| < L)
template <typename T>
__device__ __forceinline__ T wrap_pos(T x, T L) {
    if (x < 0) x += L;
    if (x >= L) x -= L;
    if (x < 0 || x >= L) x = x - L * floor(x / L);
    return x;
}

// CIC cell and weights of a position in T (the cell offsets of
// analyse/src/mitos_read.c:285-320 without the float position)
template <typename T>
__device__ __forceinline__ void cic_cell_t(T x, T icell, int N, int& i, T& d) {
    const T X = x * icell;
    T f = floor(X);
    i = (int)f;
    d = X - f;
    if (i >= N) {
        i -= N;
    }
}

// Adjoint-mode particles: moves in T, positions kept in T for the deposit
// (their float copies go to X/Y/Z), fixed-point weights 2^-shift with global
// atomics (order independent).
template <typename T>
__global__ void __launch_bounds__(256)
    k_move_deposit_s(const T* __restrict__ sx, const T* __restrict__ sy,
                     const T* __restrict__ sz, unsigned rs, int N, double boxlen, float spacing,
                     float* X, float* Y, float* Z, unsigned long long* grid, int shift) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)N * N * N;
    if (p >= n) return;
    const int qz = (int)(p % N), qy = (int)((p / N) % N), qx = (int)(p / ((long long)N * N));
    const T L = (T)boxlen, sN = (T)(N / boxlen);
    T x, y, z;
    moves_s(sx, sy, sz, rs, N, sN, spacing, qx, qy, qz, x, y, z);
    x = wrap_pos(x, L);
    y = wrap_pos(y, L);
    z = wrap_pos(z, L);
    X[p] = (float)x;
    Y[p] = (float)y;
    Z[p] = (float)z;
    int iX, iY, iZ;
    T dx, dy, dz;
    cic_cell_t(x, sN, N, iX, dx);
    cic_cell_t(y, sN, N, iY, dy);
    cic_cell_t(z, sN, N, iZ, dz);
    const T S = (T)ldexp(1.0, shift);
    const T tx = 1 - dx, ty = 1 - dy, tz = 1 - dz;
    const T w[8] = {tx * ty * tz, dx * ty * tz, tx * dy * tz, tx * ty * dz,
                    dx * dy * tz, dx * ty * dz, tx * dy * dz, dx * dy * dz};
    unsigned long long q[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        // float weights convert through 32 bits (w * 2^31 + 1/2 < 2^32), as
        // cic_weights(float) does
        if constexpr (sizeof(T) == 4)
            q[c] = (unsigned)fmaf(w[c], S, 0.5f);
        else
            q[c] = (unsigned long long)(w[c] * S + (T)0.5);
    }
    cic_global(grid, N, iX, iY, iZ, q);
}

// eight global atomic adds at base cell (iX, iY, iZ), each index in
// [-N, 2N); 32-bit cell indices (N^3 < 2^32)
__device__ __forceinline__ void cic_global_near(unsigned long long* grid, int N, int iX, int iY,
                                                int iZ, const unsigned long long w[8]) {
    const unsigned n = (unsigned)N;
    const unsigned x0 = (unsigned)wrap1(iX, N), x1 = (unsigned)wrap1((int)x0 + 1, N);
    const unsigned y0 = (unsigned)wrap1(iY, N), y1 = (unsigned)wrap1((int)y0 + 1, N);
    const unsigned z0 = (unsigned)wrap1(iZ, N), z1 = (unsigned)wrap1((int)z0 + 1, N);
    const unsigned r00 = (x0 * n + y0) * n, r10 = (x1 * n + y0) * n;
    const unsigned r01 = (x0 * n + y1) * n, r11 = (x1 * n + y1) * n;
    atomicAdd(grid + (r00 + z0), w[0]);
    atomicAdd(grid + (r10 + z0), w[1]);
    atomicAdd(grid + (r01 + z0), w[2]);
    atomicAdd(grid + (r00 + z1), w[3]);
    atomicAdd(grid + (r11 + z0), w[4]);
    atomicAdd(grid + (r10 + z1), w[5]);
    atomicAdd(grid + (r01 + z1), w[6]);
    atomicAdd(grid + (r11 + z1), w[7]);
}

// Float cell of one coordinate of a position in [0, L] (the displacement
// kernels write positions wrapped into the box): the offset in the cell is
// formed as x - i * cell with one rounding (cell = L/N is exact for the usual
// box and grid sizes, the reciprocal is not), then scaled to [0, 1]; i is
// clamped to [0, N] (N is cell 0 after the periodic offset below; the clamp
// keeps any input inside the grid).
__device__ __forceinline__ void cic_axis_f(float x, float cell, float cinv, int N, int& i,
                                           float& d) {
    int k = (int)floorf(x * cinv);
    float r = fmaf(-(float)k, cell, x);
    if (r < 0.0f) {
        k--;
        r += cell;
    } else if (r >= cell) {
        k++;
        r -= cell;
    }
    i = k < 0 ? 0 : (k > N ? N : k);
    d = fminf(r * cinv, 1.0f);
}

// Cell of a lattice particle at site (qx, qy, qz) with position (x, y, z)
// for the tiled deposit: base cell q + r with r the periodic cell offset in
// [-N/2, N/2) (so the base index lies in [-N/2, 3N/2)), and the in-cell
// offsets.
template <typename W>
__device__ __forceinline__ void cic_lattice_cell(float x, float y, float z, int N, double boxlen,
                                                 float cell, float cinv, int qx, int qy, int qz,
                                                 int& rx, int& ry, int& rz, W& fx, W& fy,
                                                 W& fz) {
    int iX, iY, iZ;
    if constexpr (sizeof(W) == 8) {
        cic_cell(x, y, z, N, boxlen, iX, iY, iZ, fx, fy, fz);
    } else {
        cic_axis_f(x, cell, cinv, N, iX, fx);
        cic_axis_f(y, cell, cinv, N, iY, fy);
        cic_axis_f(z, cell, cinv, N, iZ, fz);
    }
    rx = iX - qx;
    ry = iY - qy;
    rz = iZ - qz;
    rx = rx < -N / 2 ? rx + N : (rx >= N / 2 ? rx - N : rx);
    ry = ry < -N / 2 ? ry + N : (ry >= N / 2 ? ry - N : ry);
    rz = rz < -N / 2 ? rz + N : (rz >= N / 2 ? rz - N : rz);
}

// LDS window word: 32 bits for float weights (2^-24 units: a carry out of a
// cell, past 256 particle masses, is added to the grid as 2^32 at once), 64
// bits for double weights
template <typename W>
struct CicWin {
    typedef typename std::conditional<sizeof(W) == 4, unsigned, unsigned long long>::type type;
};

// grid index of window cell (ax, ay, az) at window origin (gx0, gy0, gz0)
// (wrapped); the window is shorter than N on x and y and than 2N on z
__device__ __forceinline__ unsigned cic_win_cell(int gx0, int gy0, int gz0, int ax, int ay,
                                                 int az, int N) {
    int gx = gx0 + ax, gy = gy0 + ay, gz = gz0 + az;
    gx = gx >= N ? gx - N : gx;
    gy = gy >= N ? gy - N : gy;
    gz = gz >= N ? gz - N : gz;
    gz = gz >= N ? gz - N : gz;
    return ((unsigned)gx * (unsigned)N + (unsigned)gy) * (unsigned)N + (unsigned)gz;
}

// carries of the 32-bit window: corner c of the cloud at window cell
// (ax, ay, az) wrapped past 2^32
__device__ __noinline__ inline void cic_win_carry(unsigned long long* grid, unsigned mask,
                                                  int gx0, int gy0, int gz0, int ax, int ay,
                                                  int az, int N) {
    // corner order of cic_weights: 000, 100, 010, 001, 110, 101, 011, 111
    const int cx[8] = {0, 1, 0, 0, 1, 1, 0, 1}, cy[8] = {0, 0, 1, 0, 1, 0, 1, 1},
              cz[8] = {0, 0, 0, 1, 0, 1, 1, 1};
    for (int c = 0; c < 8; c++)
        if (mask & (1u << c))
            atomicAdd(grid + cic_win_cell(gx0, gy0, gz0, ax + cx[c], ay + cy[c], az + cz[c], N),
                      1ull << 32);
}

// the carries of cic_win_carry appended to the carry list
struct CicLists;
__device__ __noinline__ inline void cic_list_carry(const CicLists& l, unsigned mask, int gx0,
                                                   int gy0, int gz0, int ax, int ay, int az,
                                                   int N);

// Far particles and window carries of the two-pass deposit, for the third
// pass (k_cic_far): lattice indices of particles whose cloud left their
// tile's window, and grid cells owed 2^32 by a window cell that wrapped.
// counts[0], counts[1]: entries of far and carry.
struct CicLists {
    unsigned* far;
    unsigned* carry;
    unsigned* counts;
    unsigned carry_cap;
};

__device__ __noinline__ inline void cic_list_carry(const CicLists& l, unsigned mask, int gx0,
                                                   int gy0, int gz0, int ax, int ay, int az,
                                                   int N) {
    const int cx[8] = {0, 1, 0, 0, 1, 1, 0, 1}, cy[8] = {0, 0, 1, 0, 1, 0, 1, 1},
              cz[8] = {0, 0, 0, 1, 0, 1, 1, 1};
    for (int c = 0; c < 8; c++)
        if (mask & (1u << c)) {
            const unsigned k = atomicAdd(l.counts + 1, 1u);
            if (k < l.carry_cap)  // the capacity bounds every carry a deposit can make
                l.carry[k] = cic_win_cell(gx0, gy0, gz0, ax + cx[c], ay + cy[c], az + cz[c], N);
        }
}

// CIC deposit (analyse/src/mitos_read.c:281-334) of particles stored in
// lattice order (X[(qx*N + qy)*N + qz] belongs to lattice site q).  Each
// work-group loops over PX*PY*PZ tiles of lattice sites (tile = blockIdx.x +
// k*gridDim.x, a fixed count per group, so every group exits) and deposits a
// tile into an LDS window of RX*RY*RZ cells placed at the tile's mean
// displacement; the next tile's positions are loaded while the window is
// flushed.  Window and grid sum integers, so the order of the LDS and global
// atomics does not change the result.  A particle whose cloud leaves the
// window is queued in LDS (wave-aggregated appends) and deposited with global
// atomics after the flush; the flush adds every non-zero window cell to the
// grid with one atomic and clears it.
// W = double: the reference's double weights in units of 2^-40 (FP64
// pipeline), 64-bit window; W = float: float weights in units of 2^-24 (FP32
// pipeline), 32-bit window with carries.
template <typename W, int PX, int PY, int PZ, int RX, int RY, int RZ, int NT,
          bool STREAM = false>
__global__ void __launch_bounds__(NT)
    k_cic_lattice(const float* __restrict__ X, const float* __restrict__ Y,
                  const float* __restrict__ Z, int N, double boxlen, unsigned long long* grid,
                  int ntiles, typename CicWin<W>::type* __restrict__ scratch = nullptr,
                  int* __restrict__ means = nullptr, int* __restrict__ mmax = nullptr,
                  CicLists lists = CicLists{}) {
    typedef typename CicWin<W>::type WT;
    static_assert(!STREAM || (sizeof(WT) == 4 && (RX * RY * RZ) % 4 == 0),
                  "streamed windows are 32-bit words written 16 bytes at a time");
    constexpr int P = PX * PY * PZ, K = P / NT, R = RX * RY * RZ, QCAP = 512;
    static_assert(P % NT == 0 && NT % 64 == 0, "tile must be a multiple of the work-group");
    static_assert(P <= 65536, "queue entries are 16-bit tile slots");
    // the flush wraps window cells with one (x, y) or two (z) subtractions of
    // N, so the window fits in N = 32 (the smallest grid this kernel runs on)
    static_assert(RX <= 32 && RY <= 32 && RZ <= 64, "window too large for N >= 32");
    __shared__ __align__(16) WT win[R];
    __shared__ int red[3][NT / 64];
    __shared__ unsigned short queue[QCAP];
    __shared__ int qn;
    __shared__ unsigned far_base;
    const int ntz = N / PZ, nty = N / PY;
    const float cell = (float)(boxlen / N), cinv = (float)(N / boxlen);
    const int lane = threadIdx.x & 63;
    for (int c = threadIdx.x; c < R; c += NT) win[c] = 0;
    if (threadIdx.x == 0) qn = 0;
    // positions of the group's first tile
    float px[K], py[K], pz[K];
    auto load = [&](int tile) {
        const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
#pragma unroll
        for (int i = 0; i < K; i++) {
            const int p = threadIdx.x + NT * i;
            const int lz = p % PZ, ly = (p / PZ) % PY, lx = p / (PZ * PY);
            const unsigned idx =
                ((unsigned)(tx * PX + lx) * N + (unsigned)(ty * PY + ly)) * N + tz * PZ + lz;
            px[i] = X[idx];
            py[i] = Y[idx];
            pz[i] = Z[idx];
        }
    };
    if ((int)blockIdx.x < ntiles) load(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
        // cells of the tile's particles and their mean offset
        int rxs[K], rys[K], rzs[K];
        W fx[K], fy[K], fz[K];
        int sx = 0, sy = 0, sz = 0;
#pragma unroll
        for (int i = 0; i < K; i++) {
            const int p = threadIdx.x + NT * i;
            const int lz = p % PZ, ly = (p / PZ) % PY, lx = p / (PZ * PY);
            cic_lattice_cell(px[i], py[i], pz[i], N, boxlen, cell, cinv, tx * PX + lx,
                             ty * PY + ly, tz * PZ + lz, rxs[i], rys[i], rzs[i], fx[i], fy[i],
                             fz[i]);
            sx += rxs[i];
            sy += rys[i];
            sz += rzs[i];
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
            sz += __shfl_xor(sz, o, 64);
        }
        if (lane == 0) {
            red[0][threadIdx.x >> 6] = sx;
            red[1][threadIdx.x >> 6] = sy;
            red[2][threadIdx.x >> 6] = sz;
        }
        __syncthreads();
        int mx = 0, my = 0, mz = 0;
        for (int w = 0; w < NT / 64; w++) {
            mx += red[0][w];
            my += red[1][w];
            mz += red[2][w];
        }
        const int ox = tx * PX + (int)rintf((float)mx / P) - (RX - PX - 1) / 2;
        const int oy = ty * PY + (int)rintf((float)my / P) - (RY - PY - 1) / 2;
        const int oz = tz * PZ + (int)rintf((float)mz / P) - (RZ - PZ - 1) / 2;
        const int gx0 = wrapi(ox, N), gy0 = wrapi(oy, N), gz0 = wrapi(oz, N);
#pragma unroll
        for (int i = 0; i < K; i++) {
            const int p = threadIdx.x + NT * i;
            const int lz = p % PZ, ly = (p / PZ) % PY, lx = p / (PZ * PY);
            const int qx = tx * PX + lx, qy = ty * PY + ly, qz = tz * PZ + lz;
            const int rx = rxs[i], ry = rys[i], rz = rzs[i];
            unsigned long long w[8];
            cic_weights(fx[i], fy[i], fz[i], w);
            const int ax = qx + rx - ox, ay = qy + ry - oy, az = qz + rz - oz;
            const bool in = ax >= 0 && ax + 1 < RX && ay >= 0 && ay + 1 < RY && az >= 0 &&
                            az + 1 < RZ;
            if (in) {
                WT* b = win + (ax * RY + ay) * RZ + az;
                constexpr int off[8] = {0, RY * RZ, RZ, 1, RY * RZ + RZ, RY * RZ + 1, RZ + 1,
                                        RY * RZ + RZ + 1};
                if constexpr (sizeof(WT) == 4) {
                    // a carry needs an old value of at least 2^31 (a weight
                    // is at most 2^24): test that first, the exact carries
                    // only then
                    unsigned old[8], hi = 0;
#pragma unroll
                    for (int c = 0; c < 8; c++) {
                        old[c] = atomicAdd(b + off[c], (unsigned)w[c]);
                        hi |= old[c];
                    }
                    if (hi >> 31) {
                        unsigned carry = 0;
#pragma unroll
                        for (int c = 0; c < 8; c++)
                            carry |= (old[c] + (unsigned)w[c] < old[c]) ? (1u << c) : 0u;
                        if (carry) {
                            if constexpr (STREAM)
                                cic_list_carry(lists, carry, gx0, gy0, gz0, ax, ay, az, N);
                            else
                                cic_win_carry(grid, carry, gx0, gy0, gz0, ax, ay, az, N);
                        }
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < 8; c++) atomicAdd(b + off[c], (WT)w[c]);
                }
            }
            const unsigned long long m = __ballot(!in);
            if (m) {
                const int lead = __ffsll((unsigned long long)m) - 1;
                int base = 0;
                if (lane == lead) base = atomicAdd(&qn, __popcll(m));
                base = __shfl(base, lead, 64);
                if (!in) {
                    const int slot = base + __popcll(m & ((1ull << lane) - 1));
                    if (slot < QCAP) {
                        queue[slot] = (unsigned short)(threadIdx.x + NT * i);
                    } else if constexpr (STREAM) {
                        lists.far[atomicAdd(lists.counts, 1u)] =
                            ((unsigned)qx * N + (unsigned)qy) * N + (unsigned)qz;
                    } else {
                        cic_global_near(grid, N, qx + rx, qy + ry, qz + rz, w);
                    }
                }
            }
        }
        __syncthreads();
        // the next tile's positions load while this window is flushed
        if (tile + (int)gridDim.x < ntiles) load(tile + gridDim.x);
        if constexpr (STREAM) {
            // the whole window to the tile's scratch slot (plain 16-byte
            // stores) and its mean offset to the tile table; k_cic_gather
            // adds the windows to the grid
            if (threadIdx.x == 0) {
                const int m[3] = {ox - tx * PX + (RX - PX - 1) / 2, oy - ty * PY + (RY - PY - 1) / 2,
                                  oz - tz * PZ + (RZ - PZ - 1) / 2};
                for (int a = 0; a < 3; a++) {
                    means[3 * tile + a] = m[a];
                    atomicMax(mmax + a, m[a] < 0 ? -m[a] : m[a]);
                }
            }
            uint4* dst = reinterpret_cast<uint4*>(scratch + (size_t)tile * R);
            uint4* src = reinterpret_cast<uint4*>(win);
            for (int c = threadIdx.x; c < R / 4; c += NT) {
                dst[c] = src[c];
                src[c] = make_uint4(0, 0, 0, 0);
            }
        } else {
            // window cells to the grid and cleared; (cx, cy, cz) of cell c
            // advance with c += NT
            constexpr int DZ = NT % RZ, DY = (NT / RZ) % RY, DX = NT / (RZ * RY);
            int cz = threadIdx.x % RZ, cy = (threadIdx.x / RZ) % RY,
                cx = threadIdx.x / (RZ * RY);
            for (int c = threadIdx.x; c < R; c += NT) {
                const WT v = win[c];
                if (v) {
                    win[c] = 0;
                    atomicAdd(grid + cic_win_cell(gx0, gy0, gz0, cx, cy, cz, N),
                              (unsigned long long)v);
                }
                cz += DZ;
                cy += DY;
                cx += DX;
                if (cz >= RZ) {
                    cz -= RZ;
                    cy++;
                }
                if (cy >= RY) {
                    cy -= RY;
                    cx++;
                }
            }
        }
        // queued particles: to the far list (two passes), or deposited now
        // with the same cell and weights, recomputed
        const int nq = qn < QCAP ? qn : QCAP;
        if constexpr (STREAM) {
            if (threadIdx.x == 0 && nq) far_base = atomicAdd(lists.counts, (unsigned)nq);
            __syncthreads();
            for (int j = threadIdx.x; j < nq; j += NT) {
                const int p = queue[j];
                const int lz = p % PZ, ly = (p / PZ) % PY, lx = p / (PZ * PY);
                lists.far[far_base + j] =
                    ((unsigned)(tx * PX + lx) * N + (unsigned)(ty * PY + ly)) * N + tz * PZ + lz;
            }
        }
        for (int j = threadIdx.x; j < (STREAM ? 0 : nq); j += NT) {
            const int p = queue[j];
            const int lz = p % PZ, ly = (p / PZ) % PY, lx = p / (PZ * PY);
            const int qx = tx * PX + lx, qy = ty * PY + ly, qz = tz * PZ + lz;
            const unsigned idx = ((unsigned)qx * N + (unsigned)qy) * N + qz;
            int rx, ry, rz;
            W dx, dy, dz;
            cic_lattice_cell(X[idx], Y[idx], Z[idx], N, boxlen, cell, cinv, qx, qy, qz, rx, ry,
                             rz, dx, dy, dz);
            unsigned long long w[8];
            cic_weights(dx, dy, dz, w);
            cic_global_near(grid, N, qx + rx, qy + ry, qz + rz, w);
        }
        __syncthreads();
        if (threadIdx.x == 0) qn = 0;
    }
}

// tiles of the lattice-ordered deposit kernel, 6 cells of window margin on
// each side (about 0.4% of the particles of a z = 0 amplitude field leave
// the window), 512 threads; N must be a multiple of 32.
// float: 16 x 16 x 32 sites, 28 x 28 x 44 cells of 32 bits (135 KB of LDS);
// double: 8 x 16 x 16 sites, 20 x 28 x 28 cells of 64 bits (123 KB).
template <typename W>
struct CicTile {
    static constexpr int PX = 16, PY = 16, PZ = 32, RX = 28, RY = 28, RZ = 44, NT = 512;
};
template <>
struct CicTile<double> {
    static constexpr int PX = 8, PY = 16, PZ = 16, RX = 20, RY = 28, RZ = 28, NT = 512;
};

// floor(a / b) for b > 0
__device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Second pass of the two-pass lattice deposit: one work-group per output
// tile of PX*PY*PZ cells (the cells of the lattice tile of the same index).
// Every window that can meet the tile (tile offsets bounded by the largest
// mean offset, mmax) is listed in LDS, and each thread sums its (y, z) column
// of the tile over those windows in registers; the column is then stored to
// the grid (every grid cell lies in exactly one output tile).  Each window
// cell lies in one output tile, and with R + P <= N per axis only one
// periodic image of a window meets it, so every cell is counted once; the
// sums are integers, so the order of the windows does not matter.  Eight
// waves per SIMD (64 registers, no spills): four work-groups per CU.
template <int PX, int PY, int PZ, int RX, int RY, int RZ, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8)))
    k_cic_gather(const unsigned* __restrict__ scratch, const int* __restrict__ means,
                 const int* __restrict__ mmax, int N, unsigned long long* __restrict__ grid) {
    static_assert(NT == PY * PZ, "one thread per (y, z) column of the tile");
    constexpr int R = RX * RY * RZ, PLANE = RY * RZ, MAXC = 256;
    constexpr int HX = (RX - PX - 1) / 2, HY = (RY - PY - 1) / 2, HZ = (RZ - PZ - 1) / 2;
    __shared__ int cand[MAXC][4];  // source tile and window origin relative to this tile
    __shared__ int ncand;
    const int ntz = N / PZ, nty = N / PY, ntx = N / PX;
    // z-neighbouring output tiles read the two ends of the same window rows:
    // consecutive tiles run on one XCD, so the shared lines come from its L2
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
    const int yt = threadIdx.x / PZ, zt = threadIdx.x % PZ;
    unsigned long long acc[PX];
#pragma unroll
    for (int x = 0; x < PX; x++) acc[x] = 0;
    const int MX = mmax[0], MY = mmax[1], MZ = mmax[2];
    // source tiles d with a window [d P + m - h, d P + m - h + R) that can
    // meet [0, P), |m| <= M; those that do are listed in LDS
    const int dx0 = floor_div(HX - MX - RX, PX) + 1, dx1 = floor_div(PX + HX + MX - 1, PX);
    const int dy0 = floor_div(HY - MY - RY, PY) + 1, dy1 = floor_div(PY + HY + MY - 1, PY);
    const int dz0 = floor_div(HZ - MZ - RZ, PZ) + 1, dz1 = floor_div(PZ + HZ + MZ - 1, PZ);
    const int nx = dx1 - dx0 + 1, ny = dy1 - dy0 + 1, nz = dz1 - dz0 + 1;
    const int ntot = nx * ny * nz;
    // candidates in chunks of MAXC, so the list never overflows
    for (int base = 0; base < ntot; base += MAXC) {
        if (threadIdx.x == 0) ncand = 0;
        __syncthreads();
        const int cend = base + MAXC < ntot ? base + MAXC : ntot;
        for (int c = base + threadIdx.x; c < cend; c += NT) {
            const int dz = dz0 + c % nz, dy = dy0 + (c / nz) % ny, dx = dx0 + c / (nz * ny);
            const int sx = ((tx + dx) % ntx + ntx) % ntx, sy = ((ty + dy) % nty + nty) % nty,
                      sz = ((tz + dz) % ntz + ntz) % ntz;
            const int src = (sx * nty + sy) * ntz + sz;
            const int ox = dx * PX + means[3 * src + 0] - HX;
            const int oy = dy * PY + means[3 * src + 1] - HY;
            const int oz = dz * PZ + means[3 * src + 2] - HZ;
            if (ox < PX && ox + RX > 0 && oy < PY && oy + RY > 0 && oz < PZ && oz + RZ > 0) {
                const int k = atomicAdd(&ncand, 1);
                cand[k][0] = src;
                cand[k][1] = ox;
                cand[k][2] = oy;
                cand[k][3] = oz;
            }
        }
        __syncthreads();
        const int nc = ncand;
        for (int k = 0; k < nc; k++) {
            const int src = cand[k][0], ox = cand[k][1], oy = cand[k][2], oz = cand[k][3];
            const bool in = yt >= oy && yt < oy + RY && zt >= oz && zt < oz + RZ;
            const unsigned* w = scratch + (size_t)src * R +
                                (in ? (yt - oy) * RZ + (zt - oz) : 0) - (ptrdiff_t)ox * PLANE;
            // the column's loads first (x outside the window skipped as a
            // group), then the sums
            unsigned v[PX];
#pragma unroll
            for (int x = 0; x < PX; x++) {
                v[x] = 0;
                if (x >= ox && x < ox + RX && in) v[x] = w[(ptrdiff_t)x * PLANE];
            }
#pragma unroll
            for (int x = 0; x < PX; x++) acc[x] += v[x];
        }
        __syncthreads();
    }
    const unsigned n = (unsigned)N;
    unsigned long long* g = grid + ((unsigned)(tx * PX) * n + (unsigned)(ty * PY + yt)) * n +
                            (unsigned)(tz * PZ + zt);
#pragma unroll
    for (int x = 0; x < PX; x++) g[(size_t)x * n * n] = acc[x];
}

// Third pass: far particles and window carries added to the grid.
template <typename W>
__global__ void __launch_bounds__(256)
    k_cic_far(const float* __restrict__ X, const float* __restrict__ Y,
              const float* __restrict__ Z, int N, double boxlen, unsigned long long* grid,
              CicLists lists) {
    const float cell = (float)(boxlen / N), cinv = (float)(N / boxlen);
    const unsigned nf = lists.counts[0];
    const unsigned nc = lists.counts[1] < lists.carry_cap ? lists.counts[1] : lists.carry_cap;
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < nf; j += stride) {
        const unsigned idx = lists.far[j];
        const int qz = (int)(idx % (unsigned)N), qy = (int)((idx / (unsigned)N) % (unsigned)N),
                  qx = (int)(idx / ((unsigned)N * (unsigned)N));
        int rx, ry, rz;
        W dx, dy, dz;
        cic_lattice_cell(X[idx], Y[idx], Z[idx], N, boxlen, cell, cinv, qx, qy, qz, rx, ry, rz,
                         dx, dy, dz);
        unsigned long long w[8];
        cic_weights(dx, dy, dz, w);
        cic_global_near(grid, N, qx + rx, qy + ry, qz + rz, w);
    }
    for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += stride)
        atomicAdd(grid + lists.carry[j], 1ull << 32);
}

// the tile kernel of window type W
template <typename W, bool STREAM = false, typename C = CicTile<W>>
constexpr auto cic_lattice_kernel() {
    return k_cic_lattice<W, C::PX, C::PY, C::PZ, C::RX, C::RY, C::RZ, C::NT, STREAM>;
}

static __global__ void k_zero_i32(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Workspace of the two-pass deposit (float weights, N >= 128: a window slot
// per tile, the tile mean offsets, their maxima and the list counts), in
// bytes; 0 where the two-pass deposit does not apply.
template <typename W>
inline size_t cic_stream_bytes(int N) {
    typedef CicTile<W> C;
    if (sizeof(W) != 4 || N < 128 || N % 32 != 0) return 0;
    const size_t ntiles = (size_t)(N / C::PX) * (N / C::PY) * (N / C::PZ);
    return ntiles * C::RX * C::RY * C::RZ * sizeof(unsigned) + (ntiles * 3 + 8) * sizeof(int);
}

// capacity of the carry list: a tile's window holds at most P * 2^24 units
// of mass, so it wraps at most P / 256 times
template <typename W>
inline size_t cic_carry_cap(int N) {
    typedef CicTile<W> C;
    return (size_t)N * N * N / 256 + 64;
}

// bytes of the far-particle and carry lists (every particle at most once)
template <typename W>
inline size_t cic_list_bytes(int N) {
    return ((size_t)N * N * N + cic_carry_cap<W>(N)) * sizeof(unsigned);
}

template <typename W>
inline int* cic_ws_meta(int N, void* ws) {
    typedef CicTile<W> C;
    const size_t ntiles = (size_t)(N / C::PX) * (N / C::PY) * (N / C::PZ);
    return reinterpret_cast<int*>(reinterpret_cast<unsigned*>(ws) +
                                  ntiles * C::RX * C::RY * C::RZ) +
           ntiles * 3;
}

template <typename W>
inline CicLists cic_lists(int N, void* ws, void* lists_mem) {
    CicLists l;
    l.far = reinterpret_cast<unsigned*>(lists_mem);
    l.carry = l.far + (size_t)N * N * N;
    l.counts = reinterpret_cast<unsigned*>(cic_ws_meta<W>(N, ws) + 4);
    l.carry_cap = (unsigned)cic_carry_cap<W>(N);
    return l;
}

// Launch the lattice deposit: as many work-groups as are resident at once
// (per device, weight type and mode), at most one per tile.  STREAM (two
// passes, with ws of cic_stream_bytes(N) and lists_mem of cic_list_bytes(N)):
// the windows go to ws, far particles and carries to the lists, and
// cic_gather_launch then cic_far_launch write the grid, which need not be
// cleared.  Otherwise the windows are added to a cleared grid with atomics.
template <typename W, bool STREAM>
inline int cic_lattice_launch(const float* X, const float* Y, const float* Z, int N,
                              double boxlen, unsigned long long* grid, void* ws,
                              void* lists_mem, hipStream_t s) {
    typedef CicTile<W> C;
    if (N % 32 != 0) return MITOS_ERR_UNSUPPORTED;
    if (STREAM && (!ws || !lists_mem || cic_stream_bytes<W>(N) == 0)) return MITOS_ERR_ARG;
    const int ntiles = (N / C::PX) * (N / C::PY) * (N / C::PZ);
    static std::mutex mu;
    static std::map<int, int> resident;
    int dev = 0;
    MITOS_HIP_CHECK(hipGetDevice(&dev));
    int groups = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = resident.find(dev);
        if (it == resident.end()) {
            int cus = 0, per = 0;
            MITOS_HIP_CHECK(
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            MITOS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, cic_lattice_kernel<W, STREAM>(), C::NT, 0));
            if (cus <= 0 || per <= 0) return MITOS_ERR_DEVICE;
            it = resident.emplace(dev, cus * per).first;
        }
        groups = it->second < ntiles ? it->second : ntiles;
    }
    typedef typename CicWin<W>::type WT;
    WT* scratch = nullptr;
    int* means = nullptr;
    int* meta = nullptr;
    CicLists lists{};
    if (STREAM) {
        scratch = reinterpret_cast<WT*>(ws);
        means = reinterpret_cast<int*>(scratch + (size_t)ntiles * C::RX * C::RY * C::RZ);
        meta = cic_ws_meta<W>(N, ws);
        lists = cic_lists<W>(N, ws, lists_mem);
        // mean-offset maxima (meta[0..2]) and list counts (meta[4..5]) start at 0
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, meta, 8);
    }
    hipLaunchKernelGGL((cic_lattice_kernel<W, STREAM>()), dim3(groups), dim3(C::NT), 0, s, X, Y,
                       Z, N, boxlen, grid, ntiles, scratch, means, meta, lists);
    return MITOS_OK;
}

// second pass of the two-pass deposit: windows summed into the grid
template <typename W>
inline int cic_gather_launch(int N, unsigned long long* grid, void* ws, hipStream_t s) {
    typedef CicTile<W> C;
    if (!ws || cic_stream_bytes<W>(N) == 0) return MITOS_ERR_ARG;
    const int ntiles = (N / C::PX) * (N / C::PY) * (N / C::PZ);
    const unsigned* scratch = reinterpret_cast<const unsigned*>(ws);
    const int* means =
        reinterpret_cast<const int*>(scratch + (size_t)ntiles * C::RX * C::RY * C::RZ);
    hipLaunchKernelGGL((k_cic_gather<C::PX, C::PY, C::PZ, C::RX, C::RY, C::RZ, C::NT>),
                       dim3(ntiles), dim3(C::NT), 0, s, scratch, means, cic_ws_meta<W>(N, ws), N,
                       grid);
    return MITOS_OK;
}

// third pass of the two-pass deposit: far particles and carries
template <typename W>
inline int cic_far_launch(const float* X, const float* Y, const float* Z, int N, double boxlen,
                          unsigned long long* grid, void* ws, void* lists_mem, hipStream_t s) {
    if (!ws || !lists_mem || cic_stream_bytes<W>(N) == 0) return MITOS_ERR_ARG;
    hipLaunchKernelGGL((k_cic_far<W>), dim3(1024), dim3(256), 0, s, X, Y, Z, N, boxlen, grid,
                       cic_lists<W>(N, ws, lists_mem));
    return MITOS_OK;
}

// analyse/src/mitos_read.c:363-376 for equal masses: delta = sum_w * N^3/npart - 1
template <typename T>
__global__ void k_cic_to_delta(const unsigned long long* grid, T* out, long ncell, long N,
                               long rs, double norm) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell) return;
    const long row = i / N, z = i % N;
    out[row * rs + z] = (T)((double)grid[i] * norm - 1.0);
}

}  // namespace mitos
