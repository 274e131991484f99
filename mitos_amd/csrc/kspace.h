// k-space kernels of include/fft_kernels.h evaluated on the device, and the
// closed-form spectrum (power law times BBKS transfer function).
#pragma once

#include "api_util.h"

namespace mitos {

// include/fft_kernels.h:187
__host__ __device__ __forceinline__ double sinc_d(double x) { return x == 0 ? 1. : sin(x) / x; }

// Bardeen, Bond, Kaiser & Szalay (1986), eq. G3: q = k / Gamma (k in h/Mpc)
__host__ __device__ __forceinline__ double bbks_T(double k, double Gamma) {
#pragma clang fp contract(off)
    if (k == 0) return 1.0;
    const double q = k / Gamma;
    const double a = 2.34 * q;
    const double poly = 1.0 + 3.89 * q + pow(16.1 * q, 2) + pow(5.46 * q, 3) + pow(6.71 * q, 4);
    return log(1.0 + a) / a * pow(poly, -0.25);
}

// kernel value (re, im) at one wavevector; formulas of include/fft_kernels.h
__host__ __device__ __forceinline__ cx<double> ksp_value(const KSpec& ks, double kx, double ky,
                                                         double kz, double k) {
#pragma clang fp contract(off)
    switch (ks.id) {
        case KID_CONSTANT: return mk<double>(ks.p[0], 0);               // :28-31
        case KID_INV_POISSON: return mk<double>((k > 0) ? -1.0 / k / k : 1.0, 0);  // :96-99
        case KID_DX: return mk<double>(0, kx);                         // :101-104
        case KID_DY: return mk<double>(0, ky);
        case KID_DZ: return mk<double>(0, kz);
        case KID_POWER_NO_TRANSFER: {                                  // :116-121
            if (k == 0) return mk<double>(0, 0);                       // src/primordial.c:26
            const double A_s = ks.p[0], n_s = ks.p[1], k_pivot = ks.p[2];
            return mk<double>(sqrt(A_s * pow(k / k_pivot, n_s - 1.) * k * (2. * M_PI * M_PI)),
                              0);
        }
        case KID_UNDO_WINDOW: {                                        // :198-219
            if (k == 0) return mk<double>(1.0, 0);
            const double order = ks.p[0];
            const double N = ks.p[1];
            const double boxlen = ks.p[2];
            const double W_x = sinc_d(0.5 * kx * boxlen / N);
            const double W_y = sinc_d(0.5 * ky * boxlen / N);
            const double W_z = sinc_d(0.5 * kz * boxlen / N);
            const double W = pow(W_x * W_y * W_z, order);
            return mk<double>(1.0 / W, 0);
        }
        case KID_SQRT_POWER_LAW:
            if (k == 0) return mk<double>(0, 0);
            return mk<double>(sqrt(ks.p[0] * pow(k, ks.p[1])), 0);
        case KID_BBKS:
            if (k == 0) return mk<double>(0, 0);
            return mk<double>(bbks_T(k, ks.p[0]), 0);
    }
    return mk<double>(0, 0);
}

// write[id] = read[id] * kern over the half grid (src/fft.c:257-302).  TW/TR
// may be float (the _float variants compute the product in double, as C does
// for fftwf_complex * double complex).
template <typename TW, typename TR>
__global__ void k_apply_kernel(cx<TW>* w, const cx<TR>* r, int N, long hs, long x0, long nx,
                               double boxlen, KSpec ks) {
#pragma clang fp contract(off)
    const long H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * N * H) return;
    const int z = (int)(i % H);
    const long xy = i / H;
    const int y = (int)(xy % N);
    const int x = (int)(x0 + xy / N);
    const double dk = 2 * M_PI / boxlen;
    double kx, ky, kz, k;
    wavevec(x, y, z, N, dk, kx, ky, kz, k);
    const cx<double> kv = ksp_value(ks, kx, ky, kz, k);
    const long id = ((long)(x - x0) * N + y) * hs + z;
    const cx<TR> a = r[id];
    const double re = (double)a.re, im = (double)a.im;
    w[id] = mk<TW>((TW)(re * kv.re - im * kv.im), (TW)(re * kv.im + im * kv.re));
}


// elementwise complex scaling (normalisations)
template <typename T>
__global__ void k_scale_complex(cx<T>* a, long n, double s) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cx<T> v = a[i];
    a[i] = mk<T>((T)((double)v.re * s), (T)((double)v.im * s));
}

template <typename T>
__global__ void k_div_real(T* a, long n, double d) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = (T)((double)a[i] / d);
}

}  // namespace mitos
