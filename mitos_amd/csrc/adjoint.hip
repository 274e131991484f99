// Hand-written adjoint of the IC pipeline: the gradient of a P(k) or density
// loss with respect to the white noise.
//
// The forward run is the pipeline in its stencil form (the lattice TSC
// stencils applied in k-space, k_displace_lattice_s) started from a given
// noise field.  The backward pass, in the order it runs:
//   P(k) bins        -> dL/dK(k) = 2 f^2 g_b K / (obs_b L^3)      (f: window)
//   r2c of delta     -> dL/drho(x) = c2r of the half spectrum
//   CIC deposit      -> dL/dX per particle (corner differences of the grid)
//   sequential TSC   -> scatter into the adjoints of the three stencilled
//                       displacement grids, with the position derivatives of
//                       the later moves carried to the earlier ones
//   c2r of psi       -> r2c, times the multiplicity of half-spectrum modes
//   k-space combine  -> dL/dphi1 (direct) and dL/dD (2LPT source)
//   r2c of delta2    -> dL/ddelta2(x)
//   delta2 = sum of Hessian minors (src/2lpt.c:111-112): the six phi_,ij
//                       grids are recomputed from phi1, not stored
//   phi_,ij = c2r(-k_i k_j phi1) -> six r2c, accumulated into dL/dphi1
//   Hermitian planes and the spectrum amplitude -> dL/dnoise
#include "pipe.h"

using namespace mitos;

namespace mitos {

// P(k) adjoint: K <- coef[bin] f(k)^2 K with f = s w[x] w[y] w[z] (the CIC
// window and normalisation of PkCicTransform); modes outside the bins -> 0
template <typename T>
__global__ void k_adj_pk(cx<T>* K, long hs, int N, const int* __restrict__ edges, float cinv,
                         int bins, const double* __restrict__ coef, double s,
                         const double* __restrict__ win) {
    const int H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * N * H) return;
    const int z = (int)(i % H);
    const long row = i / H;
    const int x = (int)(row / N), y = (int)(row % N);
    const int nx = freq_index(x, N), ny = freq_index(y, N);
    const long n2 = (long)nx * nx + (long)ny * ny + (long)z * z;
    cx<T>* p = K + row * hs + z;
    if (n2 == 0) {
        *p = mk<T>(0, 0);
        return;
    }
    const int b = pk_bin_of(n2, edges, bins, cinv);
    const double f = s * win[x] * win[y] * win[z];
    const double c = coef[b] * f * f;
    const cx<T> v = *p;
    *p = mk<T>((T)(v.re * c), (T)(v.im * c));
}

template <typename T>
__global__ void k_add_real(T* a, const T* __restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += b[i];
}

template <typename T>
__global__ void k_fill(T* p, long n, T v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ void wrap3(int c, int N, float invN, int o[3]) {
    const int b = wrapf(c, N, invN);
    o[0] = b == 0 ? N - 1 : b - 1;
    o[1] = b;
    o[2] = b == N - 1 ? 0 : b + 1;
}

// Per lattice particle: the moves of k_move_deposit_s recomputed in T, the
// CIC adjoint at the final position, then the adjoint of the three moves.
// sx, sy, sz: the stencilled displacement grids (halo rows of rs); grho, A*:
// dense N^3 grids.
// The y and z move adjoints scatter to neighbouring x (and y) lattice rows.
// The scatter is made order-independent (bit-reproducible gradients) in
// fixed point: pass 0 finds max|gy| and max|gz| (an integer maximum of the
// doubles' bit patterns), pass 1 adds -g w in units of max * 2^-48 to 64-bit
// integer grids (at most 2^15 maximal contributions per cell before overflow),
// and k_adj_fixed_to_real converts them.
constexpr int kAdjFixedBits = 48;

template <typename T, int PASS>
__global__ void __launch_bounds__(256)
    k_adj_particles(const T* __restrict__ sx, const T* __restrict__ sy,
                    const T* __restrict__ sz, unsigned rs, const T* __restrict__ grho, T* Ax,
                    unsigned long long* AyI, unsigned long long* AzI,
                    unsigned long long* gmax, int N, double boxlen, float spacing) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)N * N * N;
    if (p >= n) return;
    const int qz = (int)(p % N);
    const int qy = (int)((p / N) % N);
    const int qx = (int)(p / ((long long)N * N));
    const T Lb = (T)boxlen, sN = (T)(N / boxlen);
    T x1, y1, z1;
    moves_s(sx, sy, sz, rs, N, sN, spacing, qx, qy, qz, x1, y1, z1);
    // CIC (analyse/src/mitos_read.c:285-320): d rho(corner) / d position
    int iX, iY, iZ;
    T dx, dy, dz;
    cic_cell_t(wrap_pos(x1, Lb), sN, N, iX, dx);
    cic_cell_t(wrap_pos(y1, Lb), sN, N, iY, dy);
    cic_cell_t(wrap_pos(z1, Lb), sN, N, iZ, dz);
    const long x0 = wrapi(iX, N), xp = wrap1((int)x0 + 1, N);
    const long y0 = wrapi(iY, N), yp = wrap1((int)y0 + 1, N);
    const long z0 = wrapi(iZ, N), zp = wrap1((int)z0 + 1, N);
    auto g = [&](long a, long b, long c) { return grho[(a * N + b) * N + c]; };
    const T g000 = g(x0, y0, z0), g100 = g(xp, y0, z0), g010 = g(x0, yp, z0),
            g001 = g(x0, y0, zp), g110 = g(xp, yp, z0), g101 = g(xp, y0, zp),
            g011 = g(x0, yp, zp), g111 = g(xp, yp, zp);
    const T tx = 1 - dx, ty = 1 - dy, tz = 1 - dz;
    const T gX = sN * (ty * tz * (g100 - g000) + dy * tz * (g110 - g010) +
                       ty * dz * (g101 - g001) + dy * dz * (g111 - g011));
    const T gY = sN * (tx * tz * (g010 - g000) + dx * tz * (g110 - g100) +
                       tx * dz * (g011 - g001) + dx * dz * (g111 - g101));
    const T gZ = sN * (tx * ty * (g001 - g000) + dx * ty * (g101 - g100) +
                       tx * dy * (g011 - g010) + dx * dy * (g111 - g110));
    // moves: x1 = x0 - Sx(q), y1 = y0 - Fy(x1), z1 = z0 - Fz(x1, y1)
    const float invN = 1.0f / N;
    int cx, cy;
    T wx[3], dwx[3], wy[3], dwy[3];
    tsc_axis_near_d(x1 * sN, cx, wx, dwx);
    tsc_axis_near_d(y1 * sN, cy, wy, dwy);
    int ix[3], iy[3];
    wrap3(cx, N, invN, ix);
    wrap3(cy, N, invN, iy);
    const unsigned zo = MITOS_PSI_SO + qz;
    T Fy_x = 0, Fz_x = 0, Fz_y = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const unsigned ra = (unsigned)ix[a] * N;
        Fy_x += dwx[a] * sy[(ra + qy) * rs + zo];
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const T v = sz[(ra + iy[b]) * rs + zo];
            Fz_x += dwx[a] * wy[b] * v;
            Fz_y += wx[a] * dwy[b] * v;
        }
    }
    const T gz = gZ;
    const T gy = gY - gz * Fz_y * sN;
    const T gx = gX - gz * Fz_x * sN - gy * Fy_x * sN;
    if constexpr (PASS == 0) {
        // positive doubles order as their bit patterns
        atomicMax(gmax, (unsigned long long)__double_as_longlong(fabs((double)gy)));
        atomicMax(gmax + 1, (unsigned long long)__double_as_longlong(fabs((double)gz)));
        return;
    }
    const double my = __longlong_as_double((long long)gmax[0]);
    const double mz = __longlong_as_double((long long)gmax[1]);
    const double sy_ = my > 0 ? ldexp(1.0, kAdjFixedBits) / my : 0.0;
    const double sz_ = mz > 0 ? ldexp(1.0, kAdjFixedBits) / mz : 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const long ra = (long)ix[a] * N;
        atomicAdd(AyI + (ra + qy) * N + qz,
                  (unsigned long long)llrint((double)(-gy * wx[a]) * sy_));
#pragma unroll
        for (int b = 0; b < 3; b++)
            atomicAdd(AzI + (ra + iy[b]) * N + qz,
                      (unsigned long long)llrint((double)(-gz * wx[a] * wy[b]) * sz_));
    }
    Ax[p] = -gx;
}

template <typename T>
__global__ void k_adj_fixed_to_real(const unsigned long long* __restrict__ I,
                                    const unsigned long long* __restrict__ gmax, int which, T* A,
                                    long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = __longlong_as_double((long long)gmax[which]);
    A[i] = (T)((double)(long long)I[i] * ldexp(m, -kAdjFixedBits));
}

// k-space combine adjoint.  R_a = r2c of the displacement adjoints.  Forward:
// X_x = i kx Sx Sy Sz Phi, X_y = i ky Sy Sz Phi, X_z = i kz Sz Phi,
// Phi = W (f1 phi1 + f2 sD ip D), psi_a = c2r(X_a) / L^3.
// Writes G1 = dL/dphi1 (direct part) and GD = dL/dD / c_k (c_k = 2 inside the
// half spectrum, 1 on the z' = 0, N/2 planes), ready for the c2r that
// gives dL/ddelta2(x).
template <typename T>
__global__ void k_adj_combine(const cx<T>* __restrict__ Rx, const cx<T>* __restrict__ Ry,
                              const cx<T>* __restrict__ Rz, cx<T>* G1, cx<T>* GD, long hs, int N,
                              double dk, double boxvol, double sD, double f1, double f2,
                              const double* __restrict__ win, const double* __restrict__ S) {
    const int H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * N * H) return;
    const int z = (int)(i % H);
    const long row = i / H;
    const int x = (int)(row / N), y = (int)(row % N);
    const long id = row * hs + z;
    const double kx = freq_index(x, N) * dk, ky = freq_index(y, N) * dk, kz = z * dk;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double ck = (z == 0 || z == N / 2) ? 1.0 : 2.0;
    const double Sx = S[x], Sy = S[y], Sz = S[z];
    const cx<T> rx = Rx[id], ry = Ry[id], rz = Rz[id];
    // sum_a conj(i k_a S..) R_a = -i (kx SxSySz Rx + ky SySz Ry + kz Sz Rz)
    const double ax = kx * Sx * Sy * Sz, ay = ky * Sy * Sz, az = kz * Sz;
    const double sre = ax * rx.re + ay * ry.re + az * rz.re;
    const double sim = ax * rx.im + ay * ry.im + az * rz.im;
    const double c = ck / boxvol * ((k2 > 0) ? win[x] * win[y] * win[z] : 1.0);
    const double gre = sim * c, gim = -sre * c;  // -i (sre + i sim) * c
    const double ip = (k2 > 0) ? -1.0 / k2 : 1.0;
    G1[id] = mk<T>((T)(f1 * gre), (T)(f1 * gim));
    const double d = f2 * sD * ip / ck;
    GD[id] = mk<T>((T)(d * gre), (T)(d * gim));
}

// -k_i k_j of the Hessian component q (xx, yy, zz, xy, yz, xz)
__device__ __forceinline__ double hess_factor(int q, double kx, double ky, double kz) {
    switch (q) {
        case 0: return -kx * kx;
        case 1: return -ky * ky;
        case 2: return -kz * kz;
        case 3: return -kx * ky;
        case 4: return -ky * kz;
        default: return -kx * kz;
    }
}

// x-pass operator of the Hessian c2r: phi1 -> -k_i k_j phi1
struct OpHess {
    int N, q;
    double dk;
    template <typename T>
    __device__ __forceinline__ cx<T> pre(cx<T> v, int x, int y, int z) const {
        const double f = hess_factor(q, freq_index(x, N) * dk, freq_index(y, N) * dk, z * dk);
        return mk<T>((T)(v.re * f), (T)(v.im * f));
    }
    template <typename T>
    __device__ __forceinline__ cx<T> post(cx<T> v, int, int, int) const {
        return v;
    }
};

// dL/dphi_,ij from dL/ddelta2, in place over the six recomputed phi_,ij grids
template <typename T>
__global__ void k_adj_minors(T* xx, T* yy, T* zz, T* xy, T* yz, T* xz,
                             const T* __restrict__ g, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T G = g[i];
    const T a = xx[i], b = yy[i], c = zz[i];
    xx[i] = G * (b + c);
    yy[i] = G * (a + c);
    zz[i] = G * (a + b);
    xy[i] = (T)-2 * G * xy[i];
    yz[i] = (T)-2 * G * yz[i];
    xz[i] = (T)-2 * G * xz[i];
}

// G1 += c_k (-k_i k_j) R / L^3
template <typename T>
__global__ void k_adj_hess_acc(cx<T>* G1, const cx<T>* __restrict__ R, long hs, int N, int q,
                               double dk, double boxvol) {
    const int H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * N * H) return;
    const int z = (int)(i % H);
    const long row = i / H;
    const int x = (int)(row / N), y = (int)(row % N);
    const long id = row * hs + z;
    const double ck = (z == 0 || z == N / 2) ? 1.0 : 2.0;
    const double f = ck / boxvol * hess_factor(q, freq_index(x, N) * dk, freq_index(y, N) * dk,
                                               z * dk);
    const cx<T> r = R[id];
    cx<T> g = G1[id];
    g.re = (T)(g.re + f * r.re);
    g.im = (T)(g.im + f * r.im);
    G1[id] = g;
}

// adjoint of k_hermitian (in place): every lower-half plane point passes its
// gradient to its mirror (conjugated) and drops it; self-conjugate points
// keep the real part
template <typename T>
__global__ void k_adj_hermitian(cx<T>* f, long hs, int N) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)(N / 2 + 1) * N;
    if (i >= 2 * per) return;
    const int z = (i < per) ? 0 : N / 2;
    const long j = i % per;
    const int x = (int)(j / N), y = (int)(j % N);
    if ((x == 0 || x == N / 2) && y > N / 2) return;
    const int invx = (x > 0) ? N - x : 0;
    const int invy = (y > 0) ? N - y : 0;
    const int invz = (z > 0) ? N - z : 0;
    const long id = ((long)x * N + y) * hs + z;
    if (invx == x && invy == y && invz == z) {
        f[id].im = 0;
    } else {
        const long mid = ((long)invx * N + invy) * hs + z;
        const cx<T> g = f[id];
        f[mid] = f[mid] + conjg(g);
        f[id] = mk<T>(0, 0);
    }
}

// dL/dnoise = sqrt(L^3/2) amp[n^2] dL/dK, into a dense N x N x (N/2+1) array
template <typename T>
__global__ void k_adj_noise(const cx<T>* __restrict__ G, long hs, cx<T>* out, int N,
                            double boxlen, const double* __restrict__ amp) {
    const int H = N / 2 + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)N * N * H) return;
    const int z = (int)(i % H);
    const long row = i / H;
    const int x = (int)(row / N), y = (int)(row % N);
    const int nx = freq_index(x, N), ny = freq_index(y, N);
    const long n2 = (long)nx * nx + (long)ny * ny + (long)z * z;
    const double f = sqrt(boxlen * boxlen * boxlen / 2) * amp[n2];
    const cx<T> g = G[row * hs + z];
    out[i] = n2 ? mk<T>((T)(g.re * f), (T)(g.im * f)) : mk<T>(0, 0);
}

template <typename T>
static int backward(Pipe<T>* P, const mitos_ic_config& cfg, const cx<T>* noise,
                    const double* dLdP, const T* dLddelta, cx<T>* grad, hipStream_t s) {
    const int N = P->N;
    const long hs = P->hs;
    const double L = cfg.boxlen;
    const double boxvol = L * L * L;
    const double dk = 2 * M_PI / L;
    const long n3 = (long)N * N * N;
    const long nh = (long)N * N * (N / 2 + 1);
    const double sfwd = boxvol / (double)n3;
    const GridLayout dense{N, hs, (long)N};
    // forward in the stencil form: K = delta(k), psi grids, positions
    MITOS_TRY(P->enqueue(cfg, s, false, nullptr, nullptr, nullptr, noise, 1));
    T* grho = reinterpret_cast<T*>(P->b(0));
    // 1. P(k) and density adjoints -> dL/drho
    if (dLdP) {
        std::vector<long long> obs(cfg.bins);
        MITOS_HIP_CHECK(hipMemcpyAsync(obs.data(), P->pk_obs, sizeof(long long) * cfg.bins,
                                       hipMemcpyDeviceToHost, s));
        MITOS_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<double> coef(cfg.bins);
        for (int b = 0; b < cfg.bins; b++)
            coef[b] = obs[b] > 0 ? 2.0 * dLdP[b] / ((double)obs[b] * boxvol) : 0.0;
        double* dcoef = P->pk_sums;  // bins doubles of scratch (the sums are fetched)
        MITOS_HIP_CHECK(hipMemcpyAsync(dcoef, coef.data(), sizeof(double) * cfg.bins,
                                       hipMemcpyHostToDevice, s));
        int err = MITOS_OK;
        const int* edges = pk_bin_edges(N, L, cfg.bins, &err);
        if (!edges) return err;
        const double* w2 = window_table(N, L, 2);
        if (!w2) return MITOS_ERR_ALLOC;
        hipLaunchKernelGGL((k_adj_pk<T>), dim3(nblk(nh)), dim3(256), 0, s, P->K, hs, N, edges,
                           pk_bin_cinv(N, cfg.bins), cfg.bins, dcoef, sfwd, w2);
        MITOS_TRY(launch_status());
        MITOS_TRY((c2r_3d<T>(P->K, P->a(0), grho, dense, 1, 0, 0, (T)1, OpNone(), OpNone(),
                             OpNone(), s)));
        MITOS_HIP_CHECK(hipStreamSynchronize(s));  // dcoef is host-backed until here
    } else {
        hipLaunchKernelGGL((k_fill<T>), dim3(nblk(n3)), dim3(256), 0, s, grho, n3, (T)0);
        MITOS_TRY(launch_status());
    }
    if (dLddelta) {
        // delta = rho * N^3/npart - 1 with npart = N^3
        hipLaunchKernelGGL((k_add_real<T>), dim3(nblk(n3)), dim3(256), 0, s, grho, dLddelta, n3);
        MITOS_TRY(launch_status());
    }
    // 2. particles: CIC and TSC adjoints
    T* Ax = reinterpret_cast<T*>(P->b(1));
    T* Ay = reinterpret_cast<T*>(P->b(2));
    T* Az = reinterpret_cast<T*>(P->a(1));
    if (!P->adj_ws) MITOS_HIP_CHECK(hipMalloc(&P->adj_ws, sizeof(unsigned long long) *
                                                          (2 * (size_t)n3 + 2)));
    unsigned long long* AyI = P->adj_ws;
    unsigned long long* AzI = AyI + n3;
    unsigned long long* gmax = AzI + n3;
    hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(nblk(2 * n3 + 2)), dim3(256), 0, s,
                       AyI, 2 * n3 + 2, 0ULL);
    hipLaunchKernelGGL((k_adj_particles<T, 0>), dim3(nblk(n3)), dim3(256), 0, s, P->psi(0),
                       P->psi(1), P->psi(2), (unsigned)P->prs(), grho, Ax, AyI, AzI, gmax, N, L,
                       (float)L / N);
    hipLaunchKernelGGL((k_adj_particles<T, 1>), dim3(nblk(n3)), dim3(256), 0, s, P->psi(0),
                       P->psi(1), P->psi(2), (unsigned)P->prs(), grho, Ax, AyI, AzI, gmax, N, L,
                       (float)L / N);
    hipLaunchKernelGGL((k_adj_fixed_to_real<T>), dim3(nblk(n3)), dim3(256), 0, s, AyI, gmax, 0,
                       Ay, n3);
    hipLaunchKernelGGL((k_adj_fixed_to_real<T>), dim3(nblk(n3)), dim3(256), 0, s, AzI, gmax, 1,
                       Az, n3);
    MITOS_TRY(launch_status());
    // 3. back to k-space: R_a = r2c(A_a)
    cx<T>* R[3] = {P->b(3), P->b(4), P->b(5)};
    T* A3[3] = {Ax, Ay, Az};
    for (int a = 0; a < 3; a++) {
        RealRowLoader<T> ld{A3[a], (long)N, 0};
        MITOS_TRY((r2c_3d<T>(ld, R[a], dense, 1, 0, OpNone(), OpNone(), OpNone(), s)));
    }
    // 4. combine adjoint -> G1 (a1), GD (K)
    const double* win = window_table(N, L, 3);
    const double* S = smooth_table<double>(N);
    if (!win || !S) return MITOS_ERR_ALLOC;
    cx<T>* G1 = P->a(1);
    const double f2 = cfg.run2lpt ? cfg.f2 : 0.0;
    hipLaunchKernelGGL((k_adj_combine<T>), dim3(nblk(nh)), dim3(256), 0, s, R[0], R[1], R[2], G1,
                       P->K, hs, N, dk, boxvol, sfwd, cfg.f1, f2, win, S);
    MITOS_TRY(launch_status());
    if (cfg.run2lpt) {
        // 5. dL/ddelta2(x)
        T* gd2 = reinterpret_cast<T*>(P->b(0));
        MITOS_TRY((c2r_3d<T>(P->K, P->a(0), gd2, dense, 1, 0, 0, (T)1, OpNone(), OpNone(),
                             OpNone(), s)));
        // 6. recompute phi1(k) and the six phi_,ij
        const double A = amplitude_cached(cfg.n_s, cfg.gamma, cfg.sigma8);
        const double* amp = phi1_table(N, L, A, cfg.n_s, cfg.gamma);
        if (!amp) return MITOS_ERR_ALLOC;
        hipLaunchKernelGGL((k_noise_in<T>), dim3(nblk(nh)), dim3(256), 0, s, noise, P->K, hs, N,
                           L, amp);
        MITOS_TRY(launch_status());
        MITOS_TRY(launch_hermitian<T>(P->K, hs, N, 1, 0, s));
        T* ph[6] = {reinterpret_cast<T*>(P->b(1)), reinterpret_cast<T*>(P->b(2)),
                    reinterpret_cast<T*>(P->b(3)), reinterpret_cast<T*>(P->b(4)),
                    reinterpret_cast<T*>(P->b(5)), reinterpret_cast<T*>(P->a(2))};
        for (int q = 0; q < 6; q++)
            MITOS_TRY((c2r_3d<T>(P->K, P->a(0), ph[q], dense, 1, 0, 0, (T)(1.0 / boxvol),
                                 OpHess{N, q, dk}, OpNone(), OpNone(), s)));
        // 7. Hessian-minor adjoint, then back to k-space into G1
        hipLaunchKernelGGL((k_adj_minors<T>), dim3(nblk(n3)), dim3(256), 0, s, ph[0], ph[1],
                           ph[2], ph[3], ph[4], ph[5], gd2, n3);
        MITOS_TRY(launch_status());
        for (int q = 0; q < 6; q++) {
            RealRowLoader<T> ld{ph[q], (long)N, 0};
            MITOS_TRY((r2c_3d<T>(ld, P->K, dense, 1, 0, OpNone(), OpNone(), OpNone(), s)));
            hipLaunchKernelGGL((k_adj_hess_acc<T>), dim3(nblk(nh)), dim3(256), 0, s, G1, P->K, hs,
                               N, q, dk, boxvol);
            MITOS_TRY(launch_status());
        }
    }
    // 8. Hermitian planes and amplitude -> dL/dnoise
    hipLaunchKernelGGL((k_adj_hermitian<T>), dim3(nblk(2L * (N / 2 + 1) * N)), dim3(256), 0, s,
                       G1, hs, N);
    MITOS_TRY(launch_status());
    const double A = amplitude_cached(cfg.n_s, cfg.gamma, cfg.sigma8);
    const double* amp = phi1_table(N, L, A, cfg.n_s, cfg.gamma);
    if (!amp) return MITOS_ERR_ALLOC;
    hipLaunchKernelGGL((k_adj_noise<T>), dim3(nblk(nh)), dim3(256), 0, s, G1, hs, grad, N, L, amp);
    MITOS_TRY(launch_status());
    MITOS_HIP_CHECK(hipStreamSynchronize(s));
    return MITOS_OK;
}

}  // namespace mitos

// context internals shared with pipeline.hip
struct mitos_ic_context;
namespace mitos {
Pipe<float>* ctx_pipe_f(mitos_ic_context* c);
Pipe<double>* ctx_pipe_d(mitos_ic_context* c);
hipStream_t ctx_stream(mitos_ic_context* c);
// orders work issued next on s after the context's last realization
int ctx_order(mitos_ic_context* c, hipStream_t s);
// checks cfg against the context and records its bins and box size
int ctx_use_cfg(mitos_ic_context* c, const mitos_ic_config* cfg);
}  // namespace mitos

extern "C" int mitos_ic_noise_forward(mitos_ic_context* c, const struct mitos_ic_config* cfg,
                                      const void* noise, struct mitos_ic_result* res) {
    if (!c || !cfg || !noise) return set_last_error(MITOS_ERR_ARG);
    MITOS_TRY(set_last_error(ctx_use_cfg(c, cfg)));
    hipStream_t s = ctx_stream(c);
    MITOS_TRY(set_last_error(ctx_order(c, s)));
    int rc;
    if (ctx_pipe_f(c))
        rc = ctx_pipe_f(c)->enqueue(*cfg, s, false, nullptr, nullptr, nullptr,
                                    (const cx<float>*)noise, 1);
    else
        rc = ctx_pipe_d(c)->enqueue(*cfg, s, false, nullptr, nullptr, nullptr,
                                    (const cx<double>*)noise, 1);
    MITOS_TRY(set_last_error(rc));
    MITOS_HIP_CHECK(hipStreamSynchronize(s));
    if (res) MITOS_TRY(mitos_ic_fetch_pk(c, res));
    return MITOS_OK;
}

extern "C" int mitos_ic_noise_backward(mitos_ic_context* c, const struct mitos_ic_config* cfg,
                                       const void* noise, const double* dL_dP,
                                       const void* dL_ddelta, void* grad_noise) {
    if (!c || !cfg || !noise || !grad_noise || (!dL_dP && !dL_ddelta))
        return set_last_error(MITOS_ERR_ARG);
    MITOS_TRY(set_last_error(ctx_use_cfg(c, cfg)));
    hipStream_t s = ctx_stream(c);
    MITOS_TRY(set_last_error(ctx_order(c, s)));
    int rc;
    if (ctx_pipe_f(c))
        rc = backward<float>(ctx_pipe_f(c), *cfg, (const cx<float>*)noise, dL_dP,
                             (const float*)dL_ddelta, (cx<float>*)grad_noise, s);
    else
        rc = backward<double>(ctx_pipe_d(c), *cfg, (const cx<double>*)noise, dL_dP,
                              (const double*)dL_ddelta, (cx<double>*)grad_noise, s);
    return set_last_error(rc);
}
