// Particle kernels: lattice, sequential TSC displacement, fixed-point CIC.
#pragma once

#include "mitos_common.h"

namespace mitos {

// fixed-point unit of the CIC deposit: weights are summed as integers of
// 2^-40, so the total does not depend on the order of the atomic adds
#define MITOS_CIC_SHIFT 40

__host__ __device__ __forceinline__ double fwrap_d(double x, double L) {
    return fmod(fmod(x, L) + L, L);  // include/distributed_grid.h:27
}

// one axis of the TSC window (src/grids_interp.c:291-296): the three taps with
// non-zero weight, in the reference's loop order.  X is in cell units.
__device__ __forceinline__ void tsc_axis(double X, int& i0, double w[3]) {
#pragma clang fp contract(off)
    const int iX = (int)floor(X);
    const double d = X - iX;
    const int first = (d < 0.5) ? -1 : 0;  // reference loops from floor(d-1.5)
    i0 = iX + first;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const double xx = fabs(X - (iX + first + t));
        w[t] = xx < 0.5 ? (0.75 - xx * xx) : (xx < 1.5 ? 0.5 * (1.5 - xx) * (1.5 - xx) : 0);
    }
}

__device__ __forceinline__ int wrapi(int i, int N) { return ((i % N) + N) % N; }

// gridTSC_dg (src/grids_interp.c:267-309) on one rank: all taps wrap periodically.
template <typename T>
__device__ __forceinline__ double tsc_gather(const T* __restrict__ box, long rs, int N,
                                             double boxlen, double x, double y, double z) {
#pragma clang fp contract(off)
    const double X = x * N / boxlen;
    const double Y = y * N / boxlen;
    const double Z = z * N / boxlen;
    int i0, j0, k0;
    double wx[3], wy[3], wz[3];
    tsc_axis(X, i0, wx);
    tsc_axis(Y, j0, wy);
    tsc_axis(Z, k0, wz);
    int kk[3];
#pragma unroll
    for (int c = 0; c < 3; c++) kk[c] = wrapi(k0 + c, N);
    double sum = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const long ix = wrapi(i0 + a, N);
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const long iy = wrapi(j0 + b, N);
            const T* row = box + (ix * N + iy) * rs;
            const double wab = wx[a] * wy[b];
#pragma unroll
            for (int c = 0; c < 3; c++) sum += (double)row[kk[c]] * (wab * wz[c]);
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

// src/mitos.c:827-863 then :1072-1077.  LATTICE: start from the lattice
// position (M = N) instead of reading X/Y/Z.
template <typename T, bool LATTICE>
__global__ void __launch_bounds__(256)
    k_displace(const T* __restrict__ px, const T* __restrict__ py, const T* __restrict__ pz,
               long rs, int N, double boxlen, float spacing, long long npart, float* X,
               float* Y, float* Z) {
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npart) return;
    float x, y, z;
    if (LATTICE) {
        lattice_pos(p, N, spacing, x, y, z);
    } else {
        x = X[p];
        y = Y[p];
        z = Z[p];
    }
    double d = tsc_gather(px, rs, N, boxlen, (double)x, (double)y, (double)z);
    x = (float)((double)x - d);
    d = tsc_gather(py, rs, N, boxlen, (double)x, (double)y, (double)z);
    y = (float)((double)y - d);
    d = tsc_gather(pz, rs, N, boxlen, (double)x, (double)y, (double)z);
    z = (float)((double)z - d);
    X[p] = (float)fwrap_d(x, boxlen);
    Y[p] = (float)fwrap_d(y, boxlen);
    Z[p] = (float)fwrap_d(z, boxlen);
}

// analyse/src/mitos_read.c:281-334: eight weights per particle, summed as
// 2^-40 fixed point into an unpadded N^3 grid of 64-bit integers.
static __global__ void __launch_bounds__(256)
    k_cic(const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ Z,
          long long npart, int N, double boxlen, unsigned long long* grid) {
#pragma clang fp contract(off)
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= npart) return;
    const double cell = boxlen / N;
    const double Xc = fwrap_d((double)X[l], boxlen) / cell;
    const double Yc = fwrap_d((double)Y[l], boxlen) / cell;
    const double Zc = fwrap_d((double)Z[l], boxlen) / cell;
    const int iX = (int)floor(Xc), iY = (int)floor(Yc), iZ = (int)floor(Zc);
    const double dx = Xc - iX, dy = Yc - iY, dz = Zc - iZ;
    const double tx = 1.0 - dx, ty = 1.0 - dy, tz = 1.0 - dz;
    const long x0 = wrapi(iX, N), x1 = wrapi(iX + 1, N);
    const long y0 = wrapi(iY, N), y1 = wrapi(iY + 1, N);
    const long z0 = wrapi(iZ, N), z1 = wrapi(iZ + 1, N);
    const double S = (double)(1ULL << MITOS_CIC_SHIFT);
#define MITOS_DEP(a, b, c, w) \
    atomicAdd(grid + ((a) * N + (b)) * N + (c), (unsigned long long)((w) * S + 0.5))
    MITOS_DEP(x0, y0, z0, tx * ty * tz);
    MITOS_DEP(x1, y0, z0, dx * ty * tz);
    MITOS_DEP(x0, y1, z0, tx * dy * tz);
    MITOS_DEP(x0, y0, z1, tx * ty * dz);
    MITOS_DEP(x1, y1, z0, dx * dy * tz);
    MITOS_DEP(x1, y0, z1, dx * ty * dz);
    MITOS_DEP(x0, y1, z1, tx * dy * dz);
    MITOS_DEP(x1, y1, z1, dx * dy * dz);
#undef MITOS_DEP
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
