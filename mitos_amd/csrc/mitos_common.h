// Common device/host helpers for the MI355X (gfx950) Mitos IC library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <cmath>

// Error codes returned across the C ABI (same values as include/mitos_hip.h).
// 0..2 keep the reference meanings (src/fft.c:269-277, src/2lpt.c:69-72).
#ifndef MITOS_OK
#define MITOS_OK 0
#define MITOS_ERR_DIMS 1          // non-matching grid dimensions / bad sizes
#define MITOS_ERR_NOT_MOMENTUM 2  // input grid not in momentum space
#define MITOS_ERR_DEVICE 3        // HIP runtime / device failure
#define MITOS_ERR_UNSUPPORTED 4   // size or kernel callback not available on the device
#define MITOS_ERR_ARG 5           // invalid argument (null pointer, aliasing, range)
#define MITOS_ERR_ALLOC 6         // device allocation failed
#define MITOS_ERR_REAL_SPACE 7    // input grid not in configuration space
#define MITOS_ERR_TOO_LARGE 8     // input beyond the fixed scratch of an allocation-free call
#endif

#define MITOS_HIP_CHECK(expr)                                                        \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            fprintf(stderr, "mitos_hip: %s failed: %s (%s:%d)\n", #expr,             \
                    hipGetErrorString(_e), __FILE__, __LINE__);                      \
            return MITOS_ERR_DEVICE;                                                 \
        }                                                                            \
    } while (0)

#define MITOS_TRY(expr)                  \
    do {                                 \
        int _rc = (expr);                \
        if (_rc != MITOS_OK) return _rc; \
    } while (0)

namespace mitos {

// Copy and wait until the bytes are in place.  The library's kernels run on
// non-blocking streams, which are not ordered against the null stream of a
// plain hipMemcpy, and a hipMemcpy from pageable host memory or between
// device buffers may return before the copy has completed.
inline hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return e;
}

template <typename T>
struct cx {
    T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T re, T im) {
    cx<T> r;
    r.re = re;
    r.im = im;
    return r;
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
    return mk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
    return mk<T>(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> cscale(cx<T> a, T s) {
    return mk<T>(a.re * s, a.im * s);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> conjg(cx<T> a) {
    return mk<T>(a.re, -a.im);
}
// multiply by +i (DIR=+1) or -i (DIR=-1)
template <int DIR, typename T>
__host__ __device__ __forceinline__ cx<T> mul_i(cx<T> a) {
    return DIR > 0 ? mk<T>(-a.im, a.re) : mk<T>(a.im, -a.re);
}

// Vector types for 8/16-byte loads and stores of one complex value.
template <typename T>
struct vec2;
template <>
struct vec2<float> {
    typedef float2 type;
};
template <>
struct vec2<double> {
    typedef double2 type;
};

template <typename T>
__device__ __forceinline__ cx<T> ld(const cx<T>* p) {
    typename vec2<T>::type v = *reinterpret_cast<const typename vec2<T>::type*>(p);
    return mk<T>(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void st(cx<T>* p, cx<T> a) {
    typename vec2<T>::type v;
    v.x = a.re;
    v.y = a.im;
    *reinterpret_cast<typename vec2<T>::type*>(p) = v;
}

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X, each with its
// own L2.  For kernels whose neighbouring blocks share data, renumber so that
// XCD k runs the k-th contiguous eighth of the logical blocks.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    if (nb % 8u) return b;
    return (b % 8u) * (nb / 8u) + b / 8u;
}

// The (x, y) work-group of a 2D grid after the same remap of its linear
// index (x fastest): neighbouring x tiles of one row run on one XCD and share
// its L2.
__device__ __forceinline__ void xcd_block2(unsigned& bx, unsigned& by) {
    const unsigned lb = xcd_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    bx = lb % gridDim.x;
    by = lb / gridDim.x;
}

// Signed frequency index of grid index i (src/fft.c:31-33: index N/2 -> +N/2).
__host__ __device__ __forceinline__ int freq_index(int i, int N) { return (i > N / 2) ? i - N : i; }

// Layout of one realization's grids in device memory.
//   complex half grid: element (x, y, z') at (x*N + y)*hs + z', z' in [0, N/2]
//   real grid:         element (x, y, z ) at (x*N + y)*rs + z , z  in [0, N)
struct GridLayout {
    int N;
    long hs;  // complex row stride (elements), >= N/2+1
    long rs;  // real row stride (elements), >= N
};

}  // namespace mitos
