// Drop-in FFT, k-space kernel, Poisson and 2LPT entry points
// (include/fft.h:90-116, include/fft_kernels.h, include/poisson.h:29-30,
//  include/2lpt.h:27-29, include/grids_interp.h:41-48).
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/mitos_hip.h"
#include "kspace.h"

using namespace mitos;

namespace {

inline long half_elems(int N) { return (long)N * N * (N / 2 + 1); }

inline double* kern_parts(struct kernel* k) { return reinterpret_cast<double*>(&k->kern); }

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Map a reference kernel callback to its device kernel.
int kspec_from_callback(void (*compute)(struct kernel*), const void* params, int N,
                        double boxlen, KSpec* ks) {
    memset(ks, 0, sizeof(*ks));
    (void)N;
    (void)boxlen;
    if (compute == kernel_constant) {
        if (!params) return MITOS_ERR_ARG;
        ks->id = KID_CONSTANT;
        ks->p[0] = *(const double*)params;
    } else if (compute == kernel_inv_poisson) {
        ks->id = KID_INV_POISSON;
    } else if (compute == kernel_dx) {
        ks->id = KID_DX;
    } else if (compute == kernel_dy) {
        ks->id = KID_DY;
    } else if (compute == kernel_dz) {
        ks->id = KID_DZ;
    } else if (compute == kernel_power_no_transfer) {
        if (!params) return MITOS_ERR_ARG;
        const struct cosmology* c = (const struct cosmology*)params;
        ks->id = KID_POWER_NO_TRANSFER;
        ks->p[0] = c->A_s;
        ks->p[1] = c->n_s;
        ks->p[2] = c->k_pivot;
    } else if (compute == kernel_undo_Hermite_window) {
        if (!params) return MITOS_ERR_ARG;
        const struct Hermite_kern_params* h = (const struct Hermite_kern_params*)params;
        ks->id = KID_UNDO_WINDOW;
        ks->p[0] = h->order;
        ks->p[1] = h->N;
        ks->p[2] = h->boxlen;
    } else if (compute == kernel_sqrt_power_law) {
        if (!params) return MITOS_ERR_ARG;
        const struct mitos_spectrum_params* p = (const struct mitos_spectrum_params*)params;
        ks->id = KID_SQRT_POWER_LAW;
        ks->p[0] = p->A;
        ks->p[1] = p->n_s;
    } else if (compute == kernel_bbks_transfer) {
        if (!params) return MITOS_ERR_ARG;
        const struct mitos_spectrum_params* p = (const struct mitos_spectrum_params*)params;
        ks->id = KID_BBKS;
        ks->p[0] = p->Gamma;
    } else if (compute == kernel_lowpass || compute == kernel_hipass ||
               compute == kernel_gaussian || compute == kernel_gaussian_inv) {
        if (!params) return MITOS_ERR_ARG;
        ks->id = compute == kernel_lowpass   ? KID_LOWPASS
                 : compute == kernel_hipass  ? KID_HIPASS
                 : compute == kernel_gaussian ? KID_GAUSSIAN
                                              : KID_GAUSSIAN_INV;
        ks->p[0] = *(const double*)params;
    } else if (compute == kernel_tophat) {
        if (!params) return MITOS_ERR_ARG;
        ks->id = KID_TOPHAT;
        memcpy(ks->p, params, 2 * sizeof(double));
    } else if (compute == kernel_elliptic_tophat) {
        if (!params) return MITOS_ERR_ARG;
        ks->id = KID_ELLIPTIC_TOPHAT;
        memcpy(ks->p, params, 4 * sizeof(double));
    } else if (compute == kernel_translate) {
        if (!params) return MITOS_ERR_ARG;
        ks->id = KID_TRANSLATE;
        memcpy(ks->p, params, 3 * sizeof(double));
    } else {
        return MITOS_ERR_UNSUPPORTED;
    }
    return MITOS_OK;
}

// kernel_(inv_)transfer_function: the k vector and the selected function's
// two time rows go to the device, the modes are evaluated there
template <typename TW, typename TR>
int apply_transfer(cx<TW>* w, const cx<TR>* r, int N, long x0, long nx, double boxlen,
                   const struct spline_params* sp, int invert) {
    if (!sp || !sp->spline || !sp->spline->ptdat) return MITOS_ERR_ARG;
    const struct perturb_data* pt = sp->spline->ptdat;
    const int ks = pt->k_size, ts = pt->tau_size;
    if (ks < 2 || ts < 2 || !pt->k || !pt->delta || sp->index_src < 0 ||
        sp->index_src >= pt->n_functions || sp->tau_index < 0 || sp->tau_index > ts - 2)
        return MITOS_ERR_ARG;
    double* d = (double*)workspace(5, sizeof(double) * 3 * ks);
    if (!d) return MITOS_ERR_ALLOC;
    const double* rows = pt->delta + (size_t)sp->index_src * ks * ts + (size_t)ks * sp->tau_index;
    MITOS_HIP_CHECK(copy_sync(d, pt->k, sizeof(double) * ks, hipMemcpyHostToDevice));
    MITOS_HIP_CHECK(copy_sync(d + ks, rows, sizeof(double) * 2 * ks, hipMemcpyHostToDevice));
    TransferTab t{d, d + ks, d + 2 * ks, ks, sp->u_tau, invert};
    const long n = nx * N * (N / 2 + 1);
    hipLaunchKernelGGL((k_apply_transfer<TW, TR>), dim3(nblk(n)), dim3(256), 0, api_stream(), w,
                       r, N, (long)(N / 2 + 1), x0, nx, boxlen, t);
    MITOS_TRY(launch_status());
    return api_sync();
}

// Any other callback: evaluated on the host for every mode with the
// reference's protocol (src/fft.c:279-291: {kx, ky, kz, k, 0, params} in,
// kern out), a chunk of x-slabs at a time over host threads (the callback
// runs concurrently, as under the reference's OpenMP loop), copied to the
// device through two pinned buffers, and multiplied in there.
template <typename TW, typename TR>
int apply_kernel_host(cx<TW>* w, const cx<TR>* r, int N, long x0, long nx, double boxlen,
                      void (*compute)(struct kernel*), const void* params) {
    const long H = N / 2 + 1;
    const long plane = (long)N * H;
    const long bytes_slab = plane * (long)sizeof(cx<double>);
    const long per = std::max(1L, (32L << 20) / bytes_slab);  // slabs per chunk
    const size_t cbytes = (size_t)(per * bytes_slab);
    cx<double>* hbuf[2] = {nullptr, nullptr};
    cx<double>* dbuf = (cx<double>*)workspace(4, 2 * cbytes);
    if (!dbuf) return MITOS_ERR_ALLOC;
    hipEvent_t used[2] = {nullptr, nullptr};
    int rc = MITOS_OK;
    for (int b = 0; b < 2 && rc == MITOS_OK; b++) {
        if (hipHostMalloc((void**)&hbuf[b], cbytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&used[b], hipEventDisableTiming) != hipSuccess)
            rc = MITOS_ERR_ALLOC;
    }
    const double dk = 2 * M_PI / boxlen;
    const int nthreads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    for (long c = 0, it = 0; c < nx && rc == MITOS_OK; c += per, it++) {
        const long ns = std::min(per, nx - c);
        const int b = (int)(it & 1);
        if (it >= 2 && hipEventSynchronize(used[b]) != hipSuccess) {
            rc = MITOS_ERR_DEVICE;
            break;
        }
        cx<double>* hb = hbuf[b];
        auto worker = [&](int tid) {
            for (long row = tid; row < ns * N; row += nthreads) {
                const int x = (int)(x0 + c + row / N), y = (int)(row % N);
                for (int z = 0; z < H; z++) {
                    struct kernel kk;
                    wavevec(x, y, z, N, dk, kk.kx, kk.ky, kk.kz, kk.k);
                    double* kp = kern_parts(&kk);
                    kp[0] = 0;
                    kp[1] = 0;
                    kk.params = params;
                    compute(&kk);
                    hb[row * H + z] = mk<double>(kp[0], kp[1]);
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthreads; t++) pool.emplace_back(worker, t);
        worker(0);
        for (auto& th : pool) th.join();
        cx<double>* db = dbuf + (size_t)b * per * plane;
        if (hipMemcpyAsync(db, hb, (size_t)(ns * bytes_slab), hipMemcpyHostToDevice,
                           api_stream()) != hipSuccess) {
            rc = MITOS_ERR_DEVICE;
            break;
        }
        const long off = c * N * H;
        hipLaunchKernelGGL((k_apply_mult<TW, TR>), dim3(nblk(ns * plane)), dim3(256), 0,
                           api_stream(), w + off, r + off, db, N, H, ns);
        rc = launch_status();
        if (rc == MITOS_OK && hipEventRecord(used[b], api_stream()) != hipSuccess)
            rc = MITOS_ERR_DEVICE;
    }
    if (rc == MITOS_OK) rc = api_sync();
    else (void)api_sync();
    for (int b = 0; b < 2; b++) {
        if (hbuf[b]) hipHostFree(hbuf[b]);
        if (used[b]) hipEventDestroy(used[b]);
    }
    return rc;
}

template <typename TW, typename TR>
int apply_kernel(cx<TW>* w, const cx<TR>* r, int N, long x0, long nx, double boxlen,
                 void (*compute)(struct kernel*), const void* params) {
    if (!w || !r || N < 2 || !compute) return MITOS_ERR_ARG;
    if (compute == kernel_transfer_function || compute == kernel_inv_transfer_function)
        return apply_transfer<TW, TR>(w, r, N, x0, nx, boxlen,
                                      (const struct spline_params*)params,
                                      compute == kernel_inv_transfer_function);
    KSpec ks;
    const int kr = kspec_from_callback(compute, params, N, boxlen, &ks);
    if (kr == MITOS_ERR_UNSUPPORTED)
        return apply_kernel_host<TW, TR>(w, r, N, x0, nx, boxlen, compute, params);
    MITOS_TRY(kr);
    const long n = nx * N * (N / 2 + 1);
    hipLaunchKernelGGL((k_apply_kernel<TW, TR>), dim3(nblk(n)), dim3(256), 0, api_stream(), w,
                       r, N, (long)(N / 2 + 1), x0, nx, boxlen, ks);
    MITOS_TRY(launch_status());
    return api_sync();
}

template <typename T>
int scale_complex(cx<T>* a, long n, double s) {
    if (!a) return MITOS_ERR_ARG;
    hipLaunchKernelGGL((k_scale_complex<T>), dim3(nblk(n)), dim3(256), 0, api_stream(), a, n, s);
    MITOS_TRY(launch_status());
    return api_sync();
}

// unnormalised r2c of an N^3 array with row stride rs into the (N, N, N/2+1) layout
template <typename T>
int r2c_plain(int N, const T* in, long rs, cx<T>* out) {
    if (!in || !out) return MITOS_ERR_ARG;
    if (!size_supported(N)) return MITOS_ERR_UNSUPPORTED;
    GridLayout g{N, N / 2 + 1, rs};
    RealRowLoader<T> ld{in, rs, 0};
    MITOS_TRY((r2c_3d<T>(ld, out, g, 1, 0, OpNone(), OpNone(), OpNone(), api_stream())));
    return api_sync();
}

template <typename T>
int c2r_plain(int N, const cx<T>* in, T* out, long rs, double scale) {
    if (!in || !out) return MITOS_ERR_ARG;
    if (!size_supported(N)) return MITOS_ERR_UNSUPPORTED;
    GridLayout g{N, N / 2 + 1, rs};
    cx<T>* ws = (cx<T>*)workspace(0, sizeof(cx<T>) * half_elems(N));
    if (!ws) return MITOS_ERR_ALLOC;
    MITOS_TRY((c2r_3d<T>(in, ws, out, g, 1, 0, 0, (T)scale, OpNone(), OpNone(), OpNone(),
                         api_stream())));
    return api_sync();
}

// 2lpt.c:100-116: second-order source from the six Hessian grids (padding included)
__global__ void k_delta2(double* out, const double* const* h, long n) {
#pragma clang fp contract(off)
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double d_dxx = h[0][k], d_dxy = h[1][k], d_dxz = h[2][k];
    const double d_dyy = h[3][k], d_dyz = h[4][k], d_dzz = h[5][k];
    out[k] = d_dzz * d_dyy + d_dzz * d_dxx + d_dyy * d_dxx - d_dxy * d_dxy - d_dyz * d_dyz -
             d_dxz * d_dxz;
}

// src/monge_ampere.c:124-151: residual (1 + rho) - det(I + Hessian) of every
// entry (padding included), and per-block sums of resid^2 and rho^2 (tree in
// a fixed order, so the diagnostic does not depend on scheduling)
__global__ void __launch_bounds__(256) k_ma_residual(double* out, const double* const* h,
                                                     const double* rho, long n, double* part) {
#pragma clang fp contract(off)
    __shared__ double se[256], sn[256];
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0, q = 0;
    if (k < n) {
        const double d_dxx = h[0][k], d_dxy = h[1][k], d_dxz = h[2][k];
        const double d_dyy = h[3][k], d_dyz = h[4][k], d_dzz = h[5][k];
        const double M0 = 1 + d_dxx, M1 = d_dxy, M2 = d_dxz;
        const double M3 = d_dxy, M4 = 1 + d_dyy, M5 = d_dyz;
        const double M6 = d_dxz, M7 = d_dyz, M8 = 1 + d_dzz;
        // det3 of src/monge_ampere.c:38-42
        const double det = M0 * (M4 * M8 - M5 * M7) - M1 * (M3 * M8 - M5 * M6) +
                           M2 * (M3 * M7 - M4 * M6);
        const double r = rho[k];
        const double resid = (1 + r) - det;
        out[k] = resid;
        e = resid * resid;
        q = r * r;
    }
    se[threadIdx.x] = e;
    sn[threadIdx.x] = q;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sn[threadIdx.x] += sn[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = se[0];
        part[2 * blockIdx.x + 1] = sn[0];
    }
}

__global__ void k_add_inplace(double* a, const double* b, long n) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] += b[k];
}

// 2lpt.c:131-133
__global__ void k_combine(double* pot, const double* work, long n, double f1, double f2) {
#pragma clang fp contract(off)
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    pot[k] = pot[k] * f1 + f2 * work[k];
}

}  // namespace

// ---- kernel callbacks (host evaluation; addresses select device kernels) ----------
extern "C" void kernel_constant(struct kernel* the_kernel) {
    double* k = kern_parts(the_kernel);
    k[0] = *((const double*)the_kernel->params);
    k[1] = 0;
}
extern "C" void kernel_inv_poisson(struct kernel* the_kernel) {
    double* k = kern_parts(the_kernel);
    const double kk = the_kernel->k;
    k[0] = (kk > 0) ? -1.0 / kk / kk : 1.0;
    k[1] = 0;
}
extern "C" void kernel_dx(struct kernel* the_kernel) {
    double* k = kern_parts(the_kernel);
    k[0] = 0;
    k[1] = the_kernel->kx;
}
extern "C" void kernel_dy(struct kernel* the_kernel) {
    double* k = kern_parts(the_kernel);
    k[0] = 0;
    k[1] = the_kernel->ky;
}
extern "C" void kernel_dz(struct kernel* the_kernel) {
    double* k = kern_parts(the_kernel);
    k[0] = 0;
    k[1] = the_kernel->kz;
}
extern "C" void kernel_power_no_transfer(struct kernel* the_kernel) {
    KSpec ks;
    kspec_from_callback(kernel_power_no_transfer, the_kernel->params, 0, 0, &ks);
    cx<double> v = ksp_value(ks, the_kernel->kx, the_kernel->ky, the_kernel->kz, the_kernel->k);
    double* k = kern_parts(the_kernel);
    k[0] = v.re;
    k[1] = v.im;
}
extern "C" void kernel_undo_Hermite_window(struct kernel* the_kernel) {
    KSpec ks;
    kspec_from_callback(kernel_undo_Hermite_window, the_kernel->params, 0, 0, &ks);
    cx<double> v = ksp_value(ks, the_kernel->kx, the_kernel->ky, the_kernel->kz, the_kernel->k);
    double* k = kern_parts(the_kernel);
    k[0] = v.re;
    k[1] = v.im;
}
extern "C" void kernel_sqrt_power_law(struct kernel* the_kernel) {
    KSpec ks;
    kspec_from_callback(kernel_sqrt_power_law, the_kernel->params, 0, 0, &ks);
    cx<double> v = ksp_value(ks, the_kernel->kx, the_kernel->ky, the_kernel->kz, the_kernel->k);
    double* k = kern_parts(the_kernel);
    k[0] = v.re;
    k[1] = v.im;
}
extern "C" void kernel_bbks_transfer(struct kernel* the_kernel) {
    KSpec ks;
    kspec_from_callback(kernel_bbks_transfer, the_kernel->params, 0, 0, &ks);
    cx<double> v = ksp_value(ks, the_kernel->kx, the_kernel->ky, the_kernel->kz, the_kernel->k);
    double* k = kern_parts(the_kernel);
    k[0] = v.re;
    k[1] = v.im;
}

// host evaluation of a parameterised callback through its device formula
static void host_value(void (*self)(struct kernel*), struct kernel* the_kernel) {
    KSpec ks;
    kspec_from_callback(self, the_kernel->params, 0, 0, &ks);
    cx<double> v = ksp_value(ks, the_kernel->kx, the_kernel->ky, the_kernel->kz, the_kernel->k);
    double* k = kern_parts(the_kernel);
    k[0] = v.re;
    k[1] = v.im;
}
extern "C" void kernel_lowpass(struct kernel* the_kernel) { host_value(kernel_lowpass, the_kernel); }
extern "C" void kernel_hipass(struct kernel* the_kernel) { host_value(kernel_hipass, the_kernel); }
extern "C" void kernel_tophat(struct kernel* the_kernel) { host_value(kernel_tophat, the_kernel); }
extern "C" void kernel_elliptic_tophat(struct kernel* the_kernel) {
    host_value(kernel_elliptic_tophat, the_kernel);
}
extern "C" void kernel_translate(struct kernel* the_kernel) {
    host_value(kernel_translate, the_kernel);
}
extern "C" void kernel_gaussian(struct kernel* the_kernel) {
    host_value(kernel_gaussian, the_kernel);
}
extern "C" void kernel_gaussian_inv(struct kernel* the_kernel) {
    host_value(kernel_gaussian_inv, the_kernel);
}

// ---- FFTs on distributed grids (src/fft.c:219-254) ---------------------------------
extern "C" int fft_r2c_dg(struct distributed_grid* dg) {
    if (!dg || !dg->box || !dg->fbox) return set_last_error(MITOS_ERR_ARG);
    const int N = dg->N;
    if (!size_supported(N)) return set_last_error(MITOS_ERR_UNSUPPORTED);
    const double boxvol = dg->boxlen * dg->boxlen * dg->boxlen;
    const double s = boxvol / ((long long)N * N * N);
    GridLayout g{N, N / 2 + 1, N + 2};
    RealRowLoader<double> ld{dg->box, N + 2, 0};
    cx<double>* out = reinterpret_cast<cx<double>*>(dg->fbox);
    MITOS_TRY(set_last_error(r2c_3d<double>(ld, out, g, 1, 0, OpNone(), OpNone(),
                                            OpScale<double>{s}, api_stream())));
    dg->momentum_space = 1;
    return set_last_error(api_sync());
}

extern "C" int fft_c2r_dg(struct distributed_grid* dg) {
    if (!dg || !dg->box || !dg->fbox) return set_last_error(MITOS_ERR_ARG);
    const int N = dg->N;
    const double boxvol = dg->boxlen * dg->boxlen * dg->boxlen;
    MITOS_TRY(set_last_error(c2r_plain<double>(
        N, reinterpret_cast<const cx<double>*>(dg->fbox), dg->box, N + 2, 1.0 / boxvol)));
    dg->momentum_space = 0;
    return MITOS_OK;
}

extern "C" int fft_apply_kernel_dg(struct distributed_grid* dg_write,
                                   const struct distributed_grid* dg_read,
                                   void (*compute)(struct kernel* the_kernel),
                                   const void* params) {
    if (!dg_write || !dg_read) return set_last_error(MITOS_ERR_ARG);
    if (dg_read->NX != dg_write->NX || dg_read->N != dg_write->N)
        return set_last_error(MITOS_ERR_DIMS);
    if (dg_read->momentum_space != 1) return set_last_error(MITOS_ERR_NOT_MOMENTUM);
    MITOS_TRY(set_last_error(apply_kernel<double, double>(
        reinterpret_cast<cx<double>*>(dg_write->fbox),
        reinterpret_cast<const cx<double>*>(dg_read->fbox), dg_read->N, dg_read->X0,
        dg_read->NX, dg_read->boxlen, compute, params)));
    dg_write->momentum_space = 1;
    return MITOS_OK;
}

// ---- serial-array variants (src/fft.c:38-216) ----------------------------------------
extern "C" int fft_normalize_r2c(fftw_complex* arr, int N, double boxlen) {
    const double boxvol = boxlen * boxlen * boxlen;
    return set_last_error(scale_complex<double>(reinterpret_cast<cx<double>*>(arr),
                                                half_elems(N),
                                                boxvol / ((long long)N * N * N)));
}
extern "C" int fft_normalize_r2c_float(fftwf_complex* arr, int N, double boxlen) {
    const double boxvol = boxlen * boxlen * boxlen;
    return set_last_error(scale_complex<float>(reinterpret_cast<cx<float>*>(arr), half_elems(N),
                                               boxvol / ((long long)N * N * N)));
}
extern "C" int fft_normalize_r2c_float_mpi(fftwf_complex* farr, long int N, long int local_size,
                                           double boxlen) {
    const double boxvol = boxlen * boxlen * boxlen;
    return set_last_error(scale_complex<float>(reinterpret_cast<cx<float>*>(farr), local_size,
                                               boxvol / ((long long)N * N * N)));
}
extern "C" int fft_normalize_c2r(double* arr, int N, double boxlen) {
    if (!arr) return set_last_error(MITOS_ERR_ARG);
    const long n = (long)N * N * N;
    hipLaunchKernelGGL((k_div_real<double>), dim3(nblk(n)), dim3(256), 0, api_stream(), arr, n,
                       boxlen * boxlen * boxlen);
    MITOS_TRY(set_last_error(launch_status()));
    return set_last_error(api_sync());
}
extern "C" int fft_apply_kernel(fftw_complex* write, const fftw_complex* read, int N,
                                double boxlen, void (*compute)(struct kernel*),
                                const void* params) {
    return set_last_error(apply_kernel<double, double>(
        reinterpret_cast<cx<double>*>(write), reinterpret_cast<const cx<double>*>(read), N, 0,
        N, boxlen, compute, params));
}
extern "C" int fft_apply_kernel_float(fftwf_complex* write, const fftwf_complex* read, int N,
                                      double boxlen, void (*compute)(struct kernel*),
                                      const void* params) {
    return set_last_error(apply_kernel<float, float>(
        reinterpret_cast<cx<float>*>(write), reinterpret_cast<const cx<float>*>(read), N, 0, N,
        boxlen, compute, params));
}
extern "C" int fft_apply_kernel_float_mpi(fftwf_complex* write, const fftwf_complex* read,
                                          long int N, long int X0, long int NX, double boxlen,
                                          void (*compute)(struct kernel*),
                                          const void* params) {
    return set_last_error(apply_kernel<float, float>(
        reinterpret_cast<cx<float>*>(write), reinterpret_cast<const cx<float>*>(read), (int)N,
        X0, NX, boxlen, compute, params));
}

// ---- window functions (src/grids_interp.c:374-399, 430-455) ---------------------------
static int undo_window(void* farr, int N, double boxlen, int order, bool fp32) {
    struct Hermite_kern_params h;
    h.order = order;
    h.N = N;
    h.boxlen = boxlen;
    if (fp32)
        return apply_kernel<float, float>(reinterpret_cast<cx<float>*>(farr),
                                          reinterpret_cast<const cx<float>*>(farr), N, 0, N,
                                          boxlen, kernel_undo_Hermite_window, &h);
    return apply_kernel<double, double>(reinterpret_cast<cx<double>*>(farr),
                                        reinterpret_cast<const cx<double>*>(farr), N, 0, N,
                                        boxlen, kernel_undo_Hermite_window, &h);
}
extern "C" int undoCICWindow(fftw_complex* farr, int N, double boxlen) {
    return set_last_error(undo_window(farr, N, boxlen, 2, false));
}
extern "C" int undoTSCWindow(fftw_complex* farr, int N, double boxlen) {
    return set_last_error(undo_window(farr, N, boxlen, 3, false));
}
extern "C" int undoCICWindowFloat(fftwf_complex* farr, int N, double boxlen) {
    return set_last_error(undo_window(farr, N, boxlen, 2, true));
}
extern "C" int undoTSCWindowFloat(fftwf_complex* farr, int N, double boxlen) {
    return set_last_error(undo_window(farr, N, boxlen, 3, true));
}

// ---- FFTW-plan equivalents on unpadded arrays -------------------------------------------
extern "C" int mitos_fft_r2c(int N, const double* in, fftw_complex* out) {
    return set_last_error(r2c_plain<double>(N, in, N, reinterpret_cast<cx<double>*>(out)));
}
extern "C" int mitos_fft_c2r(int N, const fftw_complex* in, double* out) {
    return set_last_error(
        c2r_plain<double>(N, reinterpret_cast<const cx<double>*>(in), out, N, 1.0));
}
extern "C" int mitos_fft_r2c_float(int N, const float* in, fftwf_complex* out) {
    return set_last_error(r2c_plain<float>(N, in, N, reinterpret_cast<cx<float>*>(out)));
}
extern "C" int mitos_fft_c2r_float(int N, const fftwf_complex* in, float* out) {
    return set_last_error(
        c2r_plain<float>(N, reinterpret_cast<const cx<float>*>(in), out, N, 1.0));
}

// ---- Poisson (src/poisson.c:29-70) --------------------------------------------------------
extern "C" int solvePoisson_dg(struct distributed_grid* dg) {
    MITOS_TRY(fft_r2c_dg(dg));
    MITOS_TRY(fft_apply_kernel_dg(dg, dg, kernel_inv_poisson, NULL));
    return fft_c2r_dg(dg);
}

extern "C" int solvePoisson(double* phi, double* f, int N, double boxlen) {
    if (!phi || !f) return set_last_error(MITOS_ERR_ARG);
    cx<double>* fbox = (cx<double>*)workspace(1, sizeof(cx<double>) * half_elems(N));
    if (!fbox) return set_last_error(MITOS_ERR_ALLOC);
    fftw_complex* fb = reinterpret_cast<fftw_complex*>(fbox);
    MITOS_TRY(mitos_fft_r2c(N, f, fb));
    MITOS_TRY(fft_normalize_r2c(fb, N, boxlen));
    MITOS_TRY(fft_apply_kernel(fb, fb, N, boxlen, kernel_inv_poisson, NULL));
    MITOS_TRY(mitos_fft_c2r(N, fb, phi));
    return fft_normalize_c2r(phi, N, boxlen);
}

// ---- 2LPT (src/2lpt.c:38-143) --------------------------------------------------------------
extern "C" int solve2LPT(struct distributed_grid* potential, struct distributed_grid* density,
                         struct distributed_grid* workspace_grid, double factor1,
                         double factor2) {
    if (!potential || !density || !workspace_grid) return set_last_error(MITOS_ERR_ARG);
    const int N = density->N;
    // src/2lpt.c:54-60 (asserts in the reference)
    if (potential->NX != density->NX || density->NX != workspace_grid->NX ||
        potential->X0 != density->X0 || density->X0 != workspace_grid->X0 ||
        potential->N != density->N || density->N != workspace_grid->N)
        return set_last_error(MITOS_ERR_DIMS);
    if (potential->box == density->box || workspace_grid->box == density->box ||
        potential->box == workspace_grid->box)
        return set_last_error(MITOS_ERR_ARG);
    // src/2lpt.c:69-72 returns 1 here (fft_apply_kernel_dg returns 2 for the
    // same condition, src/fft.c:274)
    if (density->momentum_space != 1) return set_last_error(1);

    void (*const derivatives[3])(struct kernel*) = {kernel_dx, kernel_dy, kernel_dz};
    const int index_a[] = {0, 0, 0, 1, 1, 2};
    const int index_b[] = {0, 1, 2, 1, 2, 2};

    MITOS_TRY(fft_apply_kernel_dg(potential, density, kernel_inv_poisson, NULL));
    MITOS_TRY(fft_r2c_dg(density));  // src/2lpt.c:78

    const long padded = (long)N * N * (N + 2);
    double* hess[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    struct distributed_grid d;
    memset(&d, 0, sizeof(d));
    int rc = MITOS_OK;
    for (int j = 0; j < 6 && rc == MITOS_OK; j++)
        if (!(hess[j] = (double*)pool_alloc(sizeof(double) * padded))) rc = MITOS_ERR_ALLOC;
    if (rc == MITOS_OK) rc = alloc_local_grid(&d, N, density->boxlen, density->comm);
    double** hdev = nullptr;
    if (rc == MITOS_OK && hipMalloc(&hdev, sizeof(double*) * 6) != hipSuccess)
        rc = MITOS_ERR_ALLOC;
    for (int j = 0; j < 6 && rc == MITOS_OK; j++) {
        rc = fft_apply_kernel_dg(&d, potential, derivatives[index_a[j]], NULL);
        if (rc == MITOS_OK) rc = fft_apply_kernel_dg(&d, &d, derivatives[index_b[j]], NULL);
        if (rc == MITOS_OK) rc = fft_c2r_dg(&d);
        // on the library stream: the next derivative overwrites d
        if (rc == MITOS_OK &&
            hipMemcpyAsync(hess[j], d.box, sizeof(double) * padded, hipMemcpyDeviceToDevice,
                           api_stream()) != hipSuccess)
            rc = MITOS_ERR_DEVICE;
    }
    if (rc == MITOS_OK &&
        copy_sync(hdev, hess, sizeof(double*) * 6, hipMemcpyHostToDevice) != hipSuccess)
        rc = MITOS_ERR_DEVICE;
    if (rc == MITOS_OK) {
        hipLaunchKernelGGL(k_delta2, dim3(nblk(padded)), dim3(256), 0, api_stream(),
                           workspace_grid->box, (const double* const*)hdev, padded);
        rc = launch_status();
        if (rc == MITOS_OK) rc = api_sync();
    }
    free_local_grid(&d);
    if (hdev) hipFree(hdev);
    for (int j = 0; j < 6; j++)
        pool_free(hess[j]);
    if (rc != MITOS_OK) return set_last_error(rc);

    MITOS_TRY(solvePoisson_dg(workspace_grid));
    MITOS_TRY(fft_c2r_dg(potential));
    hipLaunchKernelGGL(k_combine, dim3(nblk(padded)), dim3(256), 0, api_stream(), potential->box,
                       workspace_grid->box, padded, factor1, factor2);
    MITOS_TRY(set_last_error(launch_status()));
    MITOS_TRY(set_last_error(api_sync()));
    return fft_r2c_dg(potential);
}

// ---- Monge-Ampere (src/monge_ampere.c:50-183) -----------------------------------------------
namespace {
thread_local std::vector<double> g_ma_eps;
}

extern "C" int solveMongeAmpere(struct distributed_grid* potential,
                                struct distributed_grid* density,
                                struct distributed_grid* workspace_grid, int cycles) {
    if (!potential || !density || !workspace_grid || cycles < 0)
        return set_last_error(MITOS_ERR_ARG);
    const int N = density->N;
    // src/monge_ampere.c:68-74 (asserts in the reference)
    if (potential->NX != density->NX || density->NX != workspace_grid->NX ||
        potential->X0 != density->X0 || density->X0 != workspace_grid->X0 ||
        potential->N != density->N || density->N != workspace_grid->N)
        return set_last_error(MITOS_ERR_DIMS);
    if (potential->box == density->box || workspace_grid->box == density->box ||
        potential->box == workspace_grid->box)
        return set_last_error(MITOS_ERR_ARG);
    if (density->momentum_space != 1) return set_last_error(1);  // :82-85
    g_ma_eps.clear();

    void (*const derivatives[3])(struct kernel*) = {kernel_dx, kernel_dy, kernel_dz};
    const int index_a[] = {0, 0, 0, 1, 1, 2};
    const int index_b[] = {0, 1, 2, 1, 2, 2};
    MITOS_TRY(fft_apply_kernel_dg(potential, density, kernel_inv_poisson, NULL));
    MITOS_TRY(fft_r2c_dg(density));  // :91 (as src/2lpt.c:78)
    if (cycles == 0) return MITOS_OK;

    const long padded = (long)N * N * (N + 2);
    const unsigned nb = nblk(padded);
    double* hess[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double** hdev = nullptr;
    double* part = nullptr;
    struct distributed_grid d;
    memset(&d, 0, sizeof(d));
    int rc = MITOS_OK;
    for (int j = 0; j < 6 && rc == MITOS_OK; j++)
        if (!(hess[j] = (double*)pool_alloc(sizeof(double) * padded))) rc = MITOS_ERR_ALLOC;
    if (rc == MITOS_OK && (hipMalloc(&hdev, sizeof(double*) * 6) != hipSuccess ||
                           hipMalloc(&part, sizeof(double) * 2 * nb) != hipSuccess))
        rc = MITOS_ERR_ALLOC;
    if (rc == MITOS_OK) rc = alloc_local_grid(&d, N, density->boxlen, density->comm);
    if (rc == MITOS_OK &&
        copy_sync(hdev, hess, sizeof(double*) * 6, hipMemcpyHostToDevice) != hipSuccess)
        rc = MITOS_ERR_DEVICE;
    std::vector<double> hp(2 * (size_t)nb);
    for (int it = 0; it < cycles && rc == MITOS_OK; it++) {
        // the six Hessian components (:103-117)
        for (int j = 0; j < 6 && rc == MITOS_OK; j++) {
            rc = fft_apply_kernel_dg(&d, potential, derivatives[index_a[j]], NULL);
            if (rc == MITOS_OK) rc = fft_apply_kernel_dg(&d, &d, derivatives[index_b[j]], NULL);
            if (rc == MITOS_OK) rc = fft_c2r_dg(&d);
            if (rc == MITOS_OK &&
                hipMemcpyAsync(hess[j], d.box, sizeof(double) * padded, hipMemcpyDeviceToDevice,
                               api_stream()) != hipSuccess)
                rc = MITOS_ERR_DEVICE;
        }
        if (rc != MITOS_OK) break;
        // residuals (:120-151)
        hipLaunchKernelGGL(k_ma_residual, dim3(nb), dim3(256), 0, api_stream(),
                           workspace_grid->box, (const double* const*)hdev, density->box, padded,
                           part);
        rc = launch_status();
        if (rc == MITOS_OK &&
            (hipMemcpyAsync(hp.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost,
                            api_stream()) != hipSuccess ||
             hipStreamSynchronize(api_stream()) != hipSuccess))
            rc = MITOS_ERR_DEVICE;
        if (rc != MITOS_OK) break;
        double eps = 0, norm = 0;
        for (unsigned b = 0; b < nb; b++) {
            eps += hp[2 * b];
            norm += hp[2 * b + 1];
        }
        // (:159-172) Poisson solve of the residual, added to the potential
        rc = solvePoisson_dg(workspace_grid);
        if (rc == MITOS_OK) rc = fft_c2r_dg(potential);
        if (rc == MITOS_OK) {
            hipLaunchKernelGGL(k_add_inplace, dim3(nb), dim3(256), 0, api_stream(),
                               potential->box, workspace_grid->box, padded);
            rc = launch_status();
        }
        if (rc == MITOS_OK) rc = api_sync();
        if (rc == MITOS_OK) rc = fft_r2c_dg(potential);
        // (:183-186) root-mean-square residual relative to the source
        g_ma_eps.push_back(sqrt((eps / norm) / ((double)N * N * N)));
    }
    free_local_grid(&d);
    if (hdev) hipFree(hdev);
    if (part) hipFree(part);
    for (int j = 0; j < 6; j++)
        pool_free(hess[j]);
    return set_last_error(rc);
}

extern "C" int mitos_monge_ampere_residuals(double* rms_eps, int n) {
    const int m = (int)g_ma_eps.size();
    for (int i = 0; i < n && i < m && rms_eps; i++) rms_eps[i] = g_ma_eps[i];
    return m;
}
