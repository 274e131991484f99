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
        case KID_LOWPASS: return mk<double>((k < ks.p[0]) ? 1.0 : 0.0, 0);          // :33-37
        case KID_HIPASS: return mk<double>((k > ks.p[0]) ? 1.0 : 0.0, 0);           // :39-43
        case KID_TOPHAT:                                                             // :45-51
            return mk<double>((k >= ks.p[0] && k <= ks.p[1]) ? 1.0 : 0.0, 0);
        case KID_ELLIPTIC_TOPHAT: {                                                  // :54-68
            const double rx = ks.p[0], ry = ks.p[1], rz = ks.p[2], eta2 = ks.p[3];
            const double kr = kx * rx + ky * ry + kz * rz;
            const double kk = kx * kx + ky * ky + kz * kz;
            const double rr = rx * rx + ry * ry + rz * rz;
            const double theta = kk * rr + (eta2 - 1) * kr * kr;
            return mk<double>((theta <= 4 * M_PI * M_PI) ? 1.0 : 0.0, 0);
        }
        case KID_TRANSLATE: {                                                        // :70-79
            const double ph = ks.p[0] * kx + ks.p[1] * ky + ks.p[2] * kz;
            return mk<double>(cos(ph), sin(ph));
        }
        case KID_GAUSSIAN: {                                                         // :81-86
            const double kR = k * ks.p[0];
            return mk<double>(exp(-kR * kR), 0);
        }
        case KID_GAUSSIAN_INV: {                                                     // :88-93
            const double kR = k * ks.p[0];
            return mk<double>(exp(kR * kR), 0);
        }
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


// kernel_(inv_)transfer_function (include/fft_kernels.h:131-180) on the
// device: one tabulated function at two neighbouring times, tau_index and
// tau_index + 1 (the rows lo and hi), interpolated bilinearly as
// perturbSplineInterp does (src/perturb_spline.c:141-160).
struct TransferTab {
    const double* k;   // k_size wavenumbers, increasing
    const double* lo;  // T(k_j) at tau_index
    const double* hi;  // T(k_j) at tau_index + 1
    int k_size;
    double u_tau;
    int invert;        // 1: kernel_inv_transfer_function (1/T)
};

// perturbSplineFindK (src/perturb_spline.c:100-138) returns the first i from
// its search start with k_i <= k <= k_{i+1}, i.e. the largest i with k_i < k
// (0 when k = k_0), and i = k_size - 2, u = 1 above the table.  The same
// interval by bisection; below k_0 (where the reference's search has no
// defined result) the first interval extrapolates linearly.
__host__ __device__ __forceinline__ double transfer_value(const TransferTab& t, double k) {
#pragma clang fp contract(off)
    const int n = t.k_size;
    int i;
    double u;
    if (k > t.k[n - 1]) {
        i = n - 2;
        u = 1.0;
    } else {
        int lo = 0, hi = n - 2;  // the answer lies in [lo, hi]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (t.k[mid] < k)
                lo = mid;
            else
                hi = mid - 1;
        }
        i = lo;
        const double left = t.k[i], right = t.k[i + 1];
        u = (k - left) / (right - left);
    }
    const double T11 = t.lo[i], T21 = t.lo[i + 1], T12 = t.hi[i], T22 = t.hi[i + 1];
    const double v = (1 - t.u_tau) * ((1 - u) * T11 + u * T21) +
                     t.u_tau * ((1 - u) * T12 + u * T22);
    return t.invert ? 1.0 / v : v;
}

template <typename TW, typename TR>
__global__ void k_apply_transfer(cx<TW>* w, const cx<TR>* r, int N, long hs, long x0, long nx,
                                 double boxlen, TransferTab t) {
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
    const double kv = (k == 0) ? 0.0 : transfer_value(t, k);  // DC ignored
    const long id = ((long)(x - x0) * N + y) * hs + z;
    const cx<TR> a = r[id];
    const double re = (double)a.re, im = (double)a.im;
    w[id] = mk<TW>((TW)(re * kv - im * 0.0), (TW)(re * 0.0 + im * kv));
}

// write = read * m over nx x-slabs, m the kernel values evaluated on the host
// (dense N x N x (N/2+1) per slab)
template <typename TW, typename TR>
__global__ void k_apply_mult(cx<TW>* w, const cx<TR>* r, const cx<double>* __restrict__ m,
                             int N, long hs, long nx) {
#pragma clang fp contract(off)
    const long H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * N * H) return;
    const long z = i % H;
    const long xy = i / H;
    const long id = xy * hs + z;
    const cx<TR> a = r[id];
    const cx<double> kv = m[i];
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
