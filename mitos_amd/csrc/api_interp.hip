// Random normals, grid interpolation at particle positions, particle
// velocities and the particle export layout:
//   sampleUniform / sampleNorm          src/random.c:27-49
//   gridTSC_dg                          src/grids_interp.c:180-207, 267-309
//   mitos_displace_velocity_tsc_dg      src/mitos.c:824-908, then :1072-1077
//   mitos_particles_export              src/mitos.c:1097-1107
#include <math.h>
#include <string.h>

#include <algorithm>

#include "../../include/mitos_hip.h"
#include "api_util.h"
#include "particles.h"

using namespace mitos;

namespace {

inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// access_grid (src/grids_interp.c:180-207): the local slice, else the left or
// right sliver, else 0
struct Slices {
    const double* left;
    const double* local;
    const double* right;
    int left_NX, left_X0, local_NX, local_X0, right_NX, right_X0;
};

__device__ __forceinline__ double access_slices(const Slices& s, int iX, int iY, int iZ, int N) {
    iX = wrapi(iX, N);
    iY = wrapi(iY, N);
    iZ = wrapi(iZ, N);
    const long row = (long)iY * (N + 2) + iZ;
    const long plane = (long)N * (N + 2);
    if (iX >= s.local_X0 && iX < s.local_X0 + s.local_NX)
        return s.local[(iX - s.local_X0) * plane + row];
    if (iX >= s.left_X0 && iX < s.left_X0 + s.left_NX)
        return s.left[(iX - s.left_X0) * plane + row];
    if (iX >= s.right_X0 && iX < s.right_X0 + s.right_NX)
        return s.right[(iX - s.right_X0) * plane + row];
    return 0;
}

// gridTSC_dg: the three taps of non-zero weight per axis (tsc_axis), summed
// in the reference's loop order with its rounding
__global__ void k_grid_tsc_slices(Slices s, double x, double y, double z, double boxlen, int N,
                                  double* out) {
#pragma clang fp contract(off)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double X = x * N / boxlen, Y = y * N / boxlen, Z = z * N / boxlen;
    int i0, j0, k0;
    double wx[3], wy[3], wz[3];
    tsc_axis(X, i0, wx);
    tsc_axis(Y, j0, wy);
    tsc_axis(Z, k0, wz);
    double sum = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++)
                sum += access_slices(s, i0 + a, j0 + b, k0 + c, N) * (wx[a] * wy[b] * wz[c]);
    *out = sum;
}

// src/mitos.c:827-863 (three sequential moves), :867-907 (velocities by TSC
// at the moved, not yet wrapped positions, stored as float), :1072-1077 (wrap)
__global__ void __launch_bounds__(256)
    k_displace_velocity(const double* __restrict__ px, const double* __restrict__ py,
                        const double* __restrict__ pz, const double* __restrict__ vx,
                        const double* __restrict__ vy, const double* __restrict__ vz, int N,
                        double boxlen, long long npart, float* X, float* Y, float* Z,
                        float* VX, float* VY, float* VZ) {
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npart) return;
    const unsigned rs = (unsigned)(N + 2);
    const float invN = 1.0f / N;
    float x = X[p], y = Y[p], z = Z[p];
    if (px) {
        double d = tsc_gather<double, double, false>(px, rs, N, invN, boxlen, x, y, z);
        x = (float)((double)x - d);
        d = tsc_gather<double, double, false>(py, rs, N, invN, boxlen, x, y, z);
        y = (float)((double)y - d);
        d = tsc_gather<double, double, false>(pz, rs, N, invN, boxlen, x, y, z);
        z = (float)((double)z - d);
    }
    if (vx) {
        VX[p] = (float)tsc_gather<double, double, false>(vx, rs, N, invN, boxlen, x, y, z);
        VY[p] = (float)tsc_gather<double, double, false>(vy, rs, N, invN, boxlen, x, y, z);
        VZ[p] = (float)tsc_gather<double, double, false>(vz, rs, N, invN, boxlen, x, y, z);
    }
    X[p] = (float)fwrap_fast(x, boxlen);
    Y[p] = (float)fwrap_fast(y, boxlen);
    Z[p] = (float)fwrap_fast(z, boxlen);
}

// the export arrays of src/mitos.c:1097-1107 for particles [p0, p0+n)
__global__ void k_pack(long long n, long long p0, const float* X, const float* Y, const float* Z,
                       const float* VX, const float* VY, const float* VZ, double mass,
                       long long first_id, double* coords, double* vels, double* masses,
                       long long* ids) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long p = p0 + i;
    if (coords) {
        coords[3 * i + 0] = X[p];
        coords[3 * i + 1] = Y[p];
        coords[3 * i + 2] = Z[p];
    }
    if (vels) {
        vels[3 * i + 0] = VX ? VX[p] : 0.0;
        vels[3 * i + 1] = VY ? VY[p] : 0.0;
        vels[3 * i + 2] = VZ ? VZ[p] : 0.0;
    }
    if (masses) masses[i] = mass;
    if (ids) ids[i] = p + first_id;
}

}  // namespace

// src/random.c:27-31
extern "C" double sampleUniform(rng_state* state) {
    const uint64_t A = rand_uint64(state);
    const double RM = (double)UINT64_MAX + 1;
    return ((double)A + 0.5) / RM;
}

// src/random.c:34-49: host Box-Muller with the host libm, as the reference
// evaluates it (the device noise kernels compute the correctly rounded value)
extern "C" double sampleNorm(rng_state* state) {
    const uint64_t A = rand_uint64(state);
    const uint64_t B = rand_uint64(state);
    const double RMax = (double)UINT64_MAX + 1;
    const double u = ((double)A + 0.5) / RMax;
    const double v = ((double)B + 0.5) / RMax;
    return sqrt(-2 * log(u)) * cos(2 * M_PI * v);
}

// src/grids_interp.c:267-309 at one position; the slices are device arrays
// of padded rows (N + 2 values); NaN and mitos_last_error() on a failure
extern "C" double gridTSC_dg(struct left_right_slice* lrs, double x, double y, double z,
                             double boxlen, int N) {
    if (!lrs || !lrs->local_slice || N < 2) {
        set_last_error(MITOS_ERR_ARG);
        return NAN;
    }
    double* d = (double*)workspace(7, sizeof(double));
    if (!d) {
        set_last_error(MITOS_ERR_ALLOC);
        return NAN;
    }
    const Slices s{lrs->left_slice, lrs->local_slice, lrs->right_slice, lrs->left_NX,
                   lrs->left_X0,    lrs->local_NX,    lrs->local_X0,    lrs->right_NX,
                   lrs->right_X0};
    hipLaunchKernelGGL(k_grid_tsc_slices, dim3(1), dim3(64), 0, api_stream(), s, x, y, z, boxlen,
                       N, d);
    double v = NAN;
    if (launch_status() != MITOS_OK ||
        hipMemcpyAsync(&v, d, sizeof(double), hipMemcpyDeviceToHost, api_stream()) !=
            hipSuccess ||
        hipStreamSynchronize(api_stream()) != hipSuccess) {
        set_last_error(MITOS_ERR_DEVICE);
        return NAN;
    }
    set_last_error(MITOS_OK);
    return v;
}

extern "C" int mitos_displace_velocity_tsc_dg(
    const struct distributed_grid* psi_x, const struct distributed_grid* psi_y,
    const struct distributed_grid* psi_z, const struct distributed_grid* vel_x,
    const struct distributed_grid* vel_y, const struct distributed_grid* vel_z, long long npart,
    float* X, float* Y, float* Z, float* VX, float* VY, float* VZ) {
    const bool disp = psi_x || psi_y || psi_z;
    const bool vel = vel_x || vel_y || vel_z;
    if (!X || !Y || !Z || npart < 0 || (disp && (!psi_x || !psi_y || !psi_z)) ||
        (vel && (!vel_x || !vel_y || !vel_z || !VX || !VY || !VZ)))
        return set_last_error(MITOS_ERR_ARG);
    const struct distributed_grid* g0 = disp ? psi_x : vel_x;
    if (!g0) return set_last_error(MITOS_ERR_ARG);
    const int N = g0->N;
    const struct distributed_grid* all[6] = {psi_x, psi_y, psi_z, vel_x, vel_y, vel_z};
    for (const struct distributed_grid* g : all) {
        if (!g) continue;
        if (g->N != N || g->boxlen != g0->boxlen) return set_last_error(MITOS_ERR_DIMS);
        if (g->momentum_space) return set_last_error(MITOS_ERR_REAL_SPACE);
        if (!g->box) return set_last_error(MITOS_ERR_ARG);
    }
    if (npart == 0) return MITOS_OK;
    if ((unsigned long long)(N + 2) * N * N > 0xffffffffULL)
        return set_last_error(MITOS_ERR_UNSUPPORTED);
    hipLaunchKernelGGL(k_displace_velocity, dim3(nblk(npart)), dim3(256), 0, api_stream(),
                       disp ? psi_x->box : nullptr, disp ? psi_y->box : nullptr,
                       disp ? psi_z->box : nullptr, vel ? vel_x->box : nullptr,
                       vel ? vel_y->box : nullptr, vel ? vel_z->box : nullptr, N, g0->boxlen,
                       npart, X, Y, Z, VX, VY, VZ);
    MITOS_TRY(set_last_error(launch_status()));
    return set_last_error(api_sync());
}

extern "C" int mitos_particles_export(long long npart, const float* X, const float* Y,
                                      const float* Z, const float* VX, const float* VY,
                                      const float* VZ, double mass, long long first_id,
                                      double* coords, double* vels, double* masses,
                                      long long* ids) {
    if (npart < 0 || (coords && (!X || !Y || !Z))) return set_last_error(MITOS_ERR_ARG);
    if (npart == 0) return MITOS_OK;
    const long long chunk = 1LL << 20;  // particles per staging pass (64 MiB)
    char* st = (char*)workspace(6, (size_t)chunk * 64);
    if (!st) return set_last_error(MITOS_ERR_ALLOC);
    double* dc = (double*)st;
    double* dv = dc + 3 * chunk;
    double* dm = dv + 3 * chunk;
    long long* di = (long long*)(dm + chunk);
    hipStream_t s = api_stream();
    for (long long p0 = 0; p0 < npart; p0 += chunk) {
        const long long n = std::min(chunk, npart - p0);
        hipLaunchKernelGGL(k_pack, dim3(nblk(n)), dim3(256), 0, s, n, p0, X, Y, Z, VX, VY, VZ,
                           mass, first_id, coords ? dc : nullptr, vels ? dv : nullptr,
                           masses ? dm : nullptr, ids ? di : nullptr);
        MITOS_TRY(set_last_error(launch_status()));
        if (coords)
            MITOS_HIP_CHECK(hipMemcpyAsync(coords + 3 * p0, dc, sizeof(double) * 3 * n,
                                           hipMemcpyDeviceToHost, s));
        if (vels)
            MITOS_HIP_CHECK(hipMemcpyAsync(vels + 3 * p0, dv, sizeof(double) * 3 * n,
                                           hipMemcpyDeviceToHost, s));
        if (masses)
            MITOS_HIP_CHECK(hipMemcpyAsync(masses + p0, dm, sizeof(double) * n,
                                           hipMemcpyDeviceToHost, s));
        if (ids)
            MITOS_HIP_CHECK(hipMemcpyAsync(ids + p0, di, sizeof(long long) * n,
                                           hipMemcpyDeviceToHost, s));
        MITOS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return MITOS_OK;
}
