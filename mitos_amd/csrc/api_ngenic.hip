// generate_ngeniclike_grf (src/grf_ngeniclike.c:28-203): the N-GenIC white
// noise scheme.  A RANLUX (ranlxd1) stream seeded with Seed fills an N x N
// table of column seeds in the spiral order of :71-88 (host, serial: it is
// one stream).  Every (x, y) column then reseeds its own stream and draws a
// phase and a non-zero amplitude per z < N/2 (:108-114), so the columns are
// independent: one device thread per column writes its modes, delta =
// L^1.5 sqrt(-ln ampl) at the phase (re = delta sin, im = delta cos), and the
// conjugate partners of the z = 0 plane (:157-195).  Nyquist modes, the DC
// mode and the z = N/2 plane stay zero.  The reference's outer `axes` loop
// (:90) repeats the identical clear-and-fill three times; it is done once.
#include <math.h>
#include <stdlib.h>

#include <vector>

#include "../../include/mitos_hip.h"
#include "api_util.h"
#include "ranlux.h"

using namespace mitos;

namespace {

__global__ void k_zero_c(cx<double>* a, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = mk<double>(0, 0);
}

// one (i, j) column of src/grf_ngeniclike.c:92-197 (local slab [X0, X0+NX))
__global__ void __launch_bounds__(256)
    k_ngenic_columns(cx<double>* fbox, const unsigned* __restrict__ seeds, int N, int NX, int X0,
                     double box, double fac) {
#pragma clang fp contract(off)
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (long)N * N) return;
    const int i = (int)(col / N), j = (int)(col % N);
    int ii = N - i;
    if (ii == N) ii = 0;
    const bool mine = i >= X0 && i < X0 + NX;
    const bool partner = ii >= X0 && ii < X0 + NX;
    if (!mine && !partner) return;
    const long H = N / 2 + 1;
    const double PI = M_PI;
    const long half = N / 2;  // Nsample / 2 in long arithmetic
    Ranlxd r;
    ranlxd_seed(r, seeds[col]);
    for (int k = 0; k < N / 2; k++) {
        const double phase = ranlxd_uniform(r) * 2 * PI;
        double ampl;
        do
            ampl = ranlxd_uniform(r);
        while (ampl == 0);
        if (i == N / 2 || j == N / 2 || k == N / 2) continue;
        if (i == 0 && j == 0 && k == 0) continue;
        double kvec[3];
        kvec[0] = (i < N / 2) ? i * 2 * PI / box : -(N - i) * 2 * PI / box;
        kvec[1] = (j < N / 2) ? j * 2 * PI / box : -(N - j) * 2 * PI / box;
        kvec[2] = (k < N / 2) ? k * 2 * PI / box : -(N - k) * 2 * PI / box;
        // SphereMode 0: a cube in k-space (:135-143)
        if (fabs(kvec[0]) * box / (2 * PI) > half) continue;
        if (fabs(kvec[1]) * box / (2 * PI) > half) continue;
        if (fabs(kvec[2]) * box / (2 * PI) > half) continue;
        const double p_of_k = 1.0 * -log(ampl);  // the spectrum is applied later
        const double delta = fac * sqrt(p_of_k) / 1.0;
        const double re = delta * sin(phase), im = delta * cos(phase);
        if (k > 0) {
            if (mine) fbox[((long)(i - X0) * N + j) * H + k] = mk<double>(re, im);
        } else if (i == 0) {
            if (j >= N / 2) continue;
            if (mine) {
                const int jj = N - j;
                fbox[((long)(i - X0) * N + j) * H] = mk<double>(re, im);
                fbox[((long)(i - X0) * N + jj) * H] = mk<double>(re, -im);
            }
        } else {
            if (i >= N / 2) continue;
            int jj = N - j;
            if (jj == N) jj = 0;
            if (mine) fbox[((long)(i - X0) * N + j) * H] = mk<double>(re, im);
            if (partner) fbox[((long)(ii - X0) * N + jj) * H] = mk<double>(re, -im);
        }
    }
}

}  // namespace

extern "C" int generate_ngeniclike_grf(fftw_complex* fbox, int N, int NX, int X0,
                                       long int block_size, double boxlen, long int seed) {
    (void)block_size;  // unused by the reference too
    if (!fbox || N < 2 || N % 2 || NX < 0 || X0 < 0 || X0 + NX > N || boxlen <= 0)
        return set_last_error(MITOS_ERR_ARG);
    // :59-61 (the factor sqrt(2) for the Gaussians of this scheme)
    double fac = pow(boxlen, 1.5) / sqrt(2);
    fac *= sqrt(2);
    // :63-88: the seed table, in the reference's order
    Ranlxd r;
    ranlxd_seed(r, (unsigned long)seed);
    std::vector<unsigned> tab;
    try {
        tab.assign((size_t)N * N, 0u);
    } catch (...) {
        return 4;  // :68-69
    }
    const long M = N;
    auto draw = [&]() { return (unsigned)(0x7fffffff * ranlxd_uniform(r)); };
    for (long i = 0; i < M / 2; i++) {
        for (long j = 0; j < i; j++) tab[i * M + j] = draw();
        for (long j = 0; j < i + 1; j++) tab[j * M + i] = draw();
        for (long j = 0; j < i; j++) tab[(M - 1 - i) * M + j] = draw();
        for (long j = 0; j < i + 1; j++) tab[(M - 1 - j) * M + i] = draw();
        for (long j = 0; j < i; j++) tab[i * M + (M - 1 - j)] = draw();
        for (long j = 0; j < i + 1; j++) tab[j * M + (M - 1 - i)] = draw();
        for (long j = 0; j < i; j++) tab[(M - 1 - i) * M + (M - 1 - j)] = draw();
        for (long j = 0; j < i + 1; j++) tab[(M - 1 - j) * M + (M - 1 - i)] = draw();
    }
    unsigned* dseeds = (unsigned*)workspace(8, sizeof(unsigned) * (size_t)N * N);
    if (!dseeds) return set_last_error(MITOS_ERR_ALLOC);
    MITOS_HIP_CHECK(copy_sync(dseeds, tab.data(), sizeof(unsigned) * (size_t)N * N,
                              hipMemcpyHostToDevice));
    cx<double>* f = reinterpret_cast<cx<double>*>(fbox);
    const long n = (long)NX * N * (N / 2 + 1);
    hipStream_t s = api_stream();
    hipLaunchKernelGGL(k_zero_c, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, n);
    MITOS_TRY(set_last_error(launch_status()));
    const long cols = (long)N * N;
    hipLaunchKernelGGL(k_ngenic_columns, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s,
                       f, dseeds, N, NX, X0, boxlen, fac);
    MITOS_TRY(set_last_error(launch_status()));
    return set_last_error(api_sync());
}
