/*
 * CPU oracle for the Mitos initial-condition hot path (FP64, plain C99).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline of the benchmark as the checker.  The product library never links
 * or calls it.  Built with -ffp-contract=off so every expression rounds
 * exactly as written (the reference's -march=native build may contract FMAs;
 * that choice cannot be reproduced without its binary and is documented).
 *
 * Every function names the reference file:line it restates.
 *
 * Pins (tests/test_oracle_pins.py): the RNG against golden draws of the
 * reference's own random_xorshift.h built from source; the FFTs against
 * numpy and a brute-force DFT; the kernels, Poisson and 2LPT against closed
 * forms; P(k) counts against an enumeration of the reference formula.
 * Parity unpinned against the reference itself (no reference output can be
 * produced here): the c2r of non-Hermitian Nyquist planes (FFTW leaves it
 * unspecified; numpy's convention is used) and the host-libm rounding of
 * log/cos in Box-Muller (pinned on this host's glibc only).
 */
#include "mitos_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_WRAP(i, N) ((((i) % (N)) + (N)) % (N))

static double orc_fwrap(double x, double L) { return fmod(fmod(x, L) + L, L); }

/* ------------------------------------------------------------------------ */
/* RNG: include/random_xorshift.h:7-59 (xoshiro256** seeded by splitmix64)  */
/* ------------------------------------------------------------------------ */
static inline uint64_t orc_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t orc_rng_next(orc_rng *st) {
    uint64_t *s = st->s;
    const uint64_t out = orc_rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = orc_rotl(s[3], 45);
    return out;
}

static uint64_t orc_splitmix(uint64_t *x) {
    uint64_t r = *x;
    *x = r + 0x9E3779B97f4A7C15ULL;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
    return r ^ (r >> 31);
}

orc_rng orc_rng_init(uint64_t seed) {
    orc_rng st;
    uint64_t sm = seed;
    for (int i = 0; i < 4; i++) st.s[i] = orc_splitmix(&sm);
    return st;
}

void orc_rng_draws(uint64_t seed, uint64_t skip, uint64_t n, uint64_t *out) {
    orc_rng st = orc_rng_init(seed);
    for (uint64_t i = 0; i < skip; i++) orc_rng_next(&st);
    for (uint64_t i = 0; i < n; i++) out[i] = orc_rng_next(&st);
}

/* draws offs[i] .. offs[i]+k-1 for n offsets in ascending order, in one pass */
void orc_rng_draws_at(uint64_t seed, const uint64_t *offs, uint64_t n, uint64_t k,
                      uint64_t *out) {
    orc_rng st = orc_rng_init(seed);
    uint64_t pos = 0, i = 0;
    uint64_t *ring = (uint64_t *)malloc(sizeof(uint64_t) * (k ? k : 1));
    while (i < n) {
        while (pos < offs[i]) {
            orc_rng_next(&st);
            pos++;
        }
        /* draws offs[i] .. offs[i]+k-1; later offsets may overlap this window */
        orc_rng cp = st;
        for (uint64_t j = 0; j < k; j++) ring[j] = orc_rng_next(&cp);
        for (uint64_t j = 0; j < k; j++) out[i * k + j] = ring[j];
        i++;
    }
    free(ring);
}

/* generator state after n draws from the seed's start */
void orc_rng_state_after(uint64_t seed, uint64_t n, uint64_t out[4]) {
    orc_rng st = orc_rng_init(seed);
    for (uint64_t i = 0; i < n; i++) orc_rng_next(&st);
    for (int i = 0; i < 4; i++) out[i] = st.s[i];
}

/* src/random.c:34-49: Box-Muller, cosine branch, two draws per normal */
double orc_sample_norm(orc_rng *st) {
    const uint64_t A = orc_rng_next(st);
    const uint64_t B = orc_rng_next(st);
    const double RMax = (double)UINT64_MAX + 1;
    const double u = ((double)A + 0.5) / RMax;
    const double v = ((double)B + 0.5) / RMax;
    return sqrt(-2 * log(u)) * cos(2 * M_PI * v);
}

void orc_sample_norms(uint64_t seed, uint64_t skip_draws, uint64_t n, double *out) {
    orc_rng st = orc_rng_init(seed);
    for (uint64_t i = 0; i < skip_draws; i++) orc_rng_next(&st);
    for (uint64_t i = 0; i < n; i++) out[i] = orc_sample_norm(&st);
}

/* ------------------------------------------------------------------------ */
/* grids: src/distributed_grid.c:23-58 on one rank                          */
/* ------------------------------------------------------------------------ */
int orc_grid_alloc(orc_grid *g, int N, double boxlen) {
    const size_t local = (size_t)N * N * (N / 2 + 1);
    g->N = N;
    g->boxlen = boxlen;
    g->momentum_space = 0;
    g->fbox = (double complex *)calloc(local, sizeof(double complex));
    g->box = (double *)calloc(2 * local, sizeof(double));
    return (g->fbox && g->box) ? 0 : 1;
}

void orc_grid_free(orc_grid *g) {
    free(g->fbox);
    free(g->box);
    g->fbox = NULL;
    g->box = NULL;
}

/* src/fft.c:29-35 */
void orc_wavevector(int x, int y, int z, int N, double dk, double *kx, double *ky,
                    double *kz, double *k) {
    *kx = (x > N / 2) ? (x - N) * dk : x * dk;
    *ky = (y > N / 2) ? (y - N) * dk : y * dk;
    *kz = (z > N / 2) ? (z - N) * dk : z * dk;
    *k = sqrt((*kx) * (*kx) + (*ky) * (*ky) + (*kz) * (*kz));
}

static inline size_t half_idx(int x, int y, int z, int N) {
    return ((size_t)x * N + y) * (N / 2 + 1) + z;
}

/* ------------------------------------------------------------------------ */
/* GRF: src/grf.c:26-69 and :78-175                                          */
/* ------------------------------------------------------------------------ */
int orc_generate_complex_grf(orc_grid *g, orc_rng *st) {
    const int N = g->N;
    const double boxlen = g->boxlen;
    const double boxvol = boxlen * boxlen * boxlen;
    const double factor = sqrt(boxvol / 2);
    const double dk = 2 * M_PI / boxlen;
    double kx, ky, kz, k;
    for (int x = 0; x < N; x++) {
        for (int y = 0; y < N; y++) {
            for (int z = 0; z <= N / 2; z++) {
                orc_wavevector(x, y, z, N, dk, &kx, &ky, &kz, &k);
                if (k > 0) {
                    double a = orc_sample_norm(st) * factor;
                    double b = orc_sample_norm(st) * factor;
                    g->fbox[half_idx(x, y, z, N)] = a + b * I;
                } else {
                    g->fbox[half_idx(x, y, z, N)] = 0;
                }
            }
        }
    }
    g->momentum_space = 1;
    return 0;
}

int orc_enforce_hermiticity(orc_grid *g) {
    const int N = g->N;
    double complex *plane = (double complex *)malloc((size_t)N * N * sizeof(double complex));
    if (!plane) return 1;
    for (int z = 0; z <= N / 2; z += N / 2) {
        for (int x = 0; x < N; x++)
            for (int y = 0; y < N; y++) plane[(size_t)x * N + y] = g->fbox[half_idx(x, y, z, N)];
        for (int x = 0; x < N; x++) {
            for (int y = 0; y < N; y++) {
                if (x > N / 2) continue;
                if ((x == 0 || x == N / 2) && y > N / 2) continue;
                int invx = (x > 0) ? N - x : 0;
                int invy = (y > 0) ? N - y : 0;
                int invz = (z > 0) ? N - z : 0;
                size_t id = half_idx(x, y, z, N);
                if (invx == x && invy == y && invz == z) {
                    g->fbox[id] = creal(g->fbox[id]);
                } else {
                    g->fbox[id] = conj(plane[(size_t)invx * N + invy]);
                }
            }
        }
        if (N / 2 == 0) break;
    }
    free(plane);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FFT: mixed-radix (2,3,4,5) Stockham on lines; FFTW r2c/c2r semantics      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    int nr;
    int radix[64];
    double complex *tw_fwd; /* exp(-2 pi i k / n) */
    double complex *tw_bwd;
} orc_plan;

static orc_plan *orc_plans[64];
static int orc_nplans = 0;

static orc_plan *orc_get_plan(int n) {
    for (int i = 0; i < orc_nplans; i++)
        if (orc_plans[i]->n == n) return orc_plans[i];
    orc_plan *p = (orc_plan *)calloc(1, sizeof(orc_plan));
    p->n = n;
    int m = n;
    while (m > 1) {
        int r;
        if (m % 4 == 0) r = 4;
        else if (m % 2 == 0) r = 2;
        else if (m % 3 == 0) r = 3;
        else if (m % 5 == 0) r = 5;
        else r = m; /* direct DFT of the remaining prime factor */
        p->radix[p->nr++] = r;
        m /= r;
    }
    p->tw_fwd = (double complex *)malloc(sizeof(double complex) * n);
    p->tw_bwd = (double complex *)malloc(sizeof(double complex) * n);
    for (int k = 0; k < n; k++) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        p->tw_fwd[k] = cos(a) - I * sin(a);
        p->tw_bwd[k] = cos(a) + I * sin(a);
    }
    orc_plans[orc_nplans++] = p;
    return p;
}

/* Unnormalised DFT of a contiguous line: X[k] = sum_j x[j] e^{sign 2 pi i jk/n}.
 * Result in x; work has n entries. */
static void orc_fft1d(double complex *x, double complex *work, const orc_plan *p, int sign) {
    const int n = p->n;
    const double complex *tw = (sign < 0) ? p->tw_fwd : p->tw_bwd;
    double complex *a = x, *b = work;
    int Ns = 1;
    for (int s = 0; s < p->nr; s++) {
        const int R = p->radix[s];
        const int nb = n / R;
        for (int j = 0; j < nb; j++) {
            double complex v[64];
            double complex vv[64];
            const int jm = j % Ns;
            const int step = n / (Ns * R);
            for (int r = 0; r < R; r++) {
                v[r] = a[j + r * nb];
                if (r && jm) v[r] *= tw[((long)r * jm * step) % n];
            }
            /* small DFT of length R with exact-table roots */
            for (int q = 0; q < R; q++) {
                double complex acc = 0;
                for (int r = 0; r < R; r++) acc += v[r] * tw[((long)r * q % R) * (n / R)];
                vv[q] = acc;
            }
            const int idxD = (j / Ns) * Ns * R + jm;
            for (int r = 0; r < R; r++) b[idxD + r * Ns] = vv[r];
        }
        double complex *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    if (a != x) memcpy(x, a, sizeof(double complex) * n);
}

/* threads of the FFT loops: 0 = the OpenMP default.  The reference's FFTW
 * plans are single-threaded (no fftw_init_threads, Makefile:7), so the CPU
 * baseline sets 1 here and keeps OpenMP only where the reference has it
 * (src/fft.c:139,279, src/mitos.c:844,1072). */
static int orc_fft_nthreads = 0;
static int orc_fft_team(void) {
#ifdef _OPENMP
    return orc_fft_nthreads > 0 ? orc_fft_nthreads : omp_get_max_threads();
#else
    return 1;
#endif
}
int orc_set_fft_threads(int n) {
    orc_fft_nthreads = n > 0 ? n : 0;
    return orc_fft_team();
}

/* transform every line of a complex 3D array along one axis (stride in elements) */
static void orc_fft_axis(double complex *data, int N, int H, int axis, int sign) {
    const orc_plan *p = orc_get_plan(N);
    /* axis 0: x (stride N*H), axis 1: y (stride H); lines indexed by the other two */
    const long stride = (axis == 0) ? (long)N * H : (long)H;
    #pragma omp parallel num_threads(orc_fft_team())
    {
        double complex *line = (double complex *)malloc(sizeof(double complex) * N);
        double complex *work = (double complex *)malloc(sizeof(double complex) * N);
        #pragma omp for
        for (long l = 0; l < (long)N * H; l++) {
            const long a = l / H, z = l % H;
            const long base = (axis == 0) ? a * H + z : a * (long)N * H + z;
            for (int i = 0; i < N; i++) line[i] = data[base + i * stride];
            orc_fft1d(line, work, p, sign);
            for (int i = 0; i < N; i++) data[base + i * stride] = line[i];
        }
        free(line);
        free(work);
    }
}

void orc_fft_r2c(int N, const double *in, int rs, double complex *out) {
    const int H = N / 2 + 1;
    const orc_plan *p = orc_get_plan(N);
    #pragma omp parallel num_threads(orc_fft_team())
    {
        double complex *line = (double complex *)malloc(sizeof(double complex) * N);
        double complex *work = (double complex *)malloc(sizeof(double complex) * N);
        #pragma omp for
        for (long l = 0; l < (long)N * N; l++) {
            for (int z = 0; z < N; z++) line[z] = in[l * rs + z];
            orc_fft1d(line, work, p, -1);
            for (int z = 0; z < H; z++) out[l * H + z] = line[z];
        }
        free(line);
        free(work);
    }
    orc_fft_axis(out, N, H, 1, -1);
    orc_fft_axis(out, N, H, 0, -1);
}

void orc_fft_c2r(int N, const double complex *in, double *out, int rs) {
    const int H = N / 2 + 1;
    const size_t tot = (size_t)N * N * H;
    double complex *tmp = (double complex *)malloc(sizeof(double complex) * tot);
    memcpy(tmp, in, sizeof(double complex) * tot);
    orc_fft_axis(tmp, N, H, 0, +1);
    orc_fft_axis(tmp, N, H, 1, +1);
    const orc_plan *p = orc_get_plan(N);
    #pragma omp parallel num_threads(orc_fft_team())
    {
        double complex *line = (double complex *)malloc(sizeof(double complex) * N);
        double complex *work = (double complex *)malloc(sizeof(double complex) * N);
        #pragma omp for
        for (long l = 0; l < (long)N * N; l++) {
            const double complex *h = tmp + l * H;
            /* last-axis c2r ignores the imaginary parts of the DC and Nyquist entries */
            line[0] = creal(h[0]);
            for (int k = 1; k < N / 2; k++) {
                line[k] = h[k];
                line[N - k] = conj(h[k]);
            }
            if (N > 1) line[N / 2] = creal(h[N / 2]);
            orc_fft1d(line, work, p, +1);
            for (int z = 0; z < N; z++) out[l * rs + z] = creal(line[z]);
        }
        free(line);
        free(work);
    }
    free(tmp);
}

/* src/fft.c:219-235 with normalisation :91-106 */
int orc_fft_r2c_dg(orc_grid *g) {
    const int N = g->N;
    orc_fft_r2c(N, g->box, N + 2, g->fbox);
    const double boxvol = g->boxlen * g->boxlen * g->boxlen;
    const size_t tot = (size_t)N * N * (N / 2 + 1);
    for (size_t i = 0; i < tot; i++) g->fbox[i] *= boxvol / ((long long)N * N * N);
    g->momentum_space = 1;
    return 0;
}

/* src/fft.c:238-254 with normalisation :110-125 (padding divided too) */
int orc_fft_c2r_dg(orc_grid *g) {
    const int N = g->N;
    orc_fft_c2r(N, g->fbox, g->box, N + 2);
    const double boxvol = g->boxlen * g->boxlen * g->boxlen;
    const size_t tot = (size_t)N * N * (N + 2);
    for (size_t i = 0; i < tot; i++) g->box[i] /= boxvol;
    g->momentum_space = 0;
    return 0;
}

/* src/fft.c:38-49 */
int orc_normalize_r2c(double complex *arr, int N, double boxlen) {
    const double boxvol = boxlen * boxlen * boxlen;
    const size_t tot = (size_t)N * N * (N / 2 + 1);
    for (size_t i = 0; i < tot; i++) arr[i] *= boxvol / ((long long)N * N * N);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* kernels: include/fft_kernels.h                                            */
/* ------------------------------------------------------------------------ */
static double orc_sinc(double x) { return x == 0 ? 1. : sin(x) / x; } /* :187 */

double orc_bbks_transfer(double k, double Gamma) {
    /* Bardeen, Bond, Kaiser & Szalay (1986) eq. G3 with q = k / Gamma (k in h/Mpc) */
    if (k == 0) return 1.0;
    const double q = k / Gamma;
    const double a = 2.34 * q;
    const double poly = 1.0 + 3.89 * q + pow(16.1 * q, 2) + pow(5.46 * q, 3) + pow(6.71 * q, 4);
    return log(1.0 + a) / a * pow(poly, -0.25);
}

double orc_kernel_value_re(int kid, const double *params, double kx, double ky, double kz,
                           double k, double *im) {
    *im = 0.0;
    switch (kid) {
    case ORC_K_CONSTANT: return params[0];
    case ORC_K_INV_POISSON: return (k > 0) ? -1.0 / k / k : 1.0;
    case ORC_K_DX: *im = kx; return 0.0;
    case ORC_K_DY: *im = ky; return 0.0;
    case ORC_K_DZ: *im = kz; return 0.0;
    case ORC_K_POWER_NO_TRANSFER: {
        /* src/primordial.c:25-33 */
        if (k == 0) return 0.0;
        const double A_s = params[0], n_s = params[1], k_pivot = params[2];
        return sqrt(A_s * pow(k / k_pivot, n_s - 1.) * k * (2. * M_PI * M_PI));
    }
    case ORC_K_UNDO_WINDOW: {
        if (k == 0) return 1.0;
        const int order = (int)params[0];
        const int N = (int)params[1];
        const double boxlen = params[2];
        double W_x = orc_sinc(0.5 * kx * boxlen / N);
        double W_y = orc_sinc(0.5 * ky * boxlen / N);
        double W_z = orc_sinc(0.5 * kz * boxlen / N);
        double W = pow(W_x * W_y * W_z, order);
        return 1.0 / W;
    }
    case ORC_K_SQRT_POWER_LAW:
        if (k == 0) return 0.0;
        return sqrt(params[0] * pow(k, params[1]));
    case ORC_K_BBKS:
        if (k == 0) return 0.0;
        return orc_bbks_transfer(k, params[0]);
    /* include/fft_kernels.h:33-93 */
    case ORC_K_LOWPASS: return (k < params[0]) ? 1.0 : 0.0;
    case ORC_K_HIPASS: return (k > params[0]) ? 1.0 : 0.0;
    case ORC_K_TOPHAT: return (k >= params[0] && k <= params[1]) ? 1.0 : 0.0;
    case ORC_K_ELLIPTIC_TOPHAT: {
        const double kr = kx * params[0] + ky * params[1] + kz * params[2];
        const double kk = kx * kx + ky * ky + kz * kz;
        const double rr = params[0] * params[0] + params[1] * params[1] + params[2] * params[2];
        const double theta = kk * rr + (params[3] - 1) * kr * kr;
        return (theta <= 4 * M_PI * M_PI) ? 1.0 : 0.0;
    }
    case ORC_K_TRANSLATE: {
        const double complex e = cexp(I * (params[0] * kx + params[1] * ky + params[2] * kz));
        *im = cimag(e);
        return creal(e);
    }
    case ORC_K_GAUSSIAN: {
        const double kR = k * params[0];
        return exp(-kR * kR);
    }
    case ORC_K_GAUSSIAN_INV: {
        const double kR = k * params[0];
        return exp(kR * kR);
    }
    }
    return 0.0;
}

static void orc_apply_kernel_arr(double complex *w, const double complex *r, int N,
                                 double boxlen, int kid, const double *params) {
    const double dk = 2 * M_PI / boxlen;
    #pragma omp parallel for
    for (int x = 0; x < N; x++) {
        for (int y = 0; y < N; y++) {
            for (int z = 0; z <= N / 2; z++) {
                double kx, ky, kz, k, im;
                orc_wavevector(x, y, z, N, dk, &kx, &ky, &kz, &k);
                double re = orc_kernel_value_re(kid, params, kx, ky, kz, k, &im);
                const size_t id = half_idx(x, y, z, N);
                const double a = creal(r[id]), b = cimag(r[id]);
                /* complex product written out (same rounding as C99 '*') */
                w[id] = (a * re - b * im) + I * (a * im + b * re);
            }
        }
    }
}

int orc_apply_kernel_dg(orc_grid *w, const orc_grid *r, int kid, const double *params) {
    if (r->N != w->N) return 1;
    if (r->momentum_space != 1) return 2;
    orc_apply_kernel_arr(w->fbox, r->fbox, r->N, r->boxlen, kid, params);
    w->momentum_space = 1;
    return 0;
}

int orc_apply_kernel(double complex *w, const double complex *r, int N, double boxlen, int kid,
                     const double *params) {
    orc_apply_kernel_arr(w, r, N, boxlen, kid, params);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* RANLUX ranlxd1 (GSL gsl_rng_ranlxd1, used by src/grf_ngeniclike.c) and   */
/* the N-GenIC white noise                                                    */
/* ------------------------------------------------------------------------ */
/* Restated from the published algorithm (Luescher 1994, the 48-bit double
 * variant; GSL's luxury level 1 = 202 steps per 12 outputs, its 31-bit
 * shift-register seeding).  GSL is absent here and its version is not
 * pinned (Makefile:9): parity unpinned. */
static void orc_ranlxd_seed(orc_ranlxd *s, unsigned long seed) {
    const double one_bit = 1.0 / 281474976710656.0;
    if (seed == 0) seed = 1;
    int i = (int)(seed & 0xFFFFFFFFUL);
    int xbit[31];
    for (int k = 0; k < 31; ++k) {
        xbit[k] = i % 2;
        i /= 2;
    }
    int ibit = 0, jbit = 18;
    for (int k = 0; k < 12; ++k) {
        double x = 0;
        for (int l = 1; l <= 48; ++l) {
            const double y = (double)((xbit[ibit] + 1) % 2);
            x += x + y;
            xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
            ibit = (ibit + 1) % 31;
            jbit = (jbit + 1) % 31;
        }
        s->x[k] = one_bit * x;
    }
    s->carry = 0;
    s->ir = 11;
    s->jr = 7;
    s->ir_old = 0;
    s->pr = 202;
}

static void orc_ranlxd_refill(orc_ranlxd *s) {
    const double one_bit = 1.0 / 281474976710656.0;
    for (int k = 0; k < s->pr; ++k) {
        double y = s->x[s->jr] - s->x[s->ir];
        y -= s->carry;
        if (y < 0) {
            s->carry = one_bit;
            y += 1;
        } else {
            s->carry = 0;
        }
        s->x[s->ir] = y;
        s->ir = (s->ir + 1) % 12;
        s->jr = (s->jr + 1) % 12;
    }
    s->ir_old = s->ir;
}

static double orc_ranlxd_uniform(orc_ranlxd *s) {
    s->ir = (s->ir + 1) % 12;
    if (s->ir == s->ir_old) orc_ranlxd_refill(s);
    return s->x[s->ir];
}

void orc_ranlxd_draws(unsigned long seed, long n, double *out) {
    orc_ranlxd s;
    orc_ranlxd_seed(&s, seed);
    for (long i = 0; i < n; i++) out[i] = orc_ranlxd_uniform(&s);
}

/* src/grf_ngeniclike.c:28-203, loop by loop (one rank: NX = N, X0 = 0) */
int orc_generate_ngeniclike_grf(double complex *Cdata, int N, double boxlen, long seed) {
    const long Nmesh = N, Nsample = N;
    const double PI = M_PI, Box = boxlen;
    double fac = pow(Box, 1.5) / sqrt(2);
    fac *= sqrt(2);
    orc_ranlxd rg;
    orc_ranlxd_seed(&rg, (unsigned long)seed);
    unsigned int *seedtable = (unsigned int *)malloc(Nmesh * Nmesh * sizeof(unsigned int));
    if (!seedtable) return 4;
    for (long i = 0; i < Nmesh / 2; i++) {
        for (long j = 0; j < i; j++)
            seedtable[i * Nmesh + j] = 0x7fffffff * orc_ranlxd_uniform(&rg);
        for (long j = 0; j < i + 1; j++)
            seedtable[j * Nmesh + i] = 0x7fffffff * orc_ranlxd_uniform(&rg);
        for (long j = 0; j < i; j++)
            seedtable[(Nmesh - 1 - i) * Nmesh + j] = 0x7fffffff * orc_ranlxd_uniform(&rg);
        for (long j = 0; j < i + 1; j++)
            seedtable[(Nmesh - 1 - j) * Nmesh + i] = 0x7fffffff * orc_ranlxd_uniform(&rg);
        for (long j = 0; j < i; j++)
            seedtable[i * Nmesh + (Nmesh - 1 - j)] = 0x7fffffff * orc_ranlxd_uniform(&rg);
        for (long j = 0; j < i + 1; j++)
            seedtable[j * Nmesh + (Nmesh - 1 - i)] = 0x7fffffff * orc_ranlxd_uniform(&rg);
        for (long j = 0; j < i; j++)
            seedtable[(Nmesh - 1 - i) * Nmesh + (Nmesh - 1 - j)] =
                0x7fffffff * orc_ranlxd_uniform(&rg);
        for (long j = 0; j < i + 1; j++)
            seedtable[(Nmesh - 1 - j) * Nmesh + (Nmesh - 1 - i)] =
                0x7fffffff * orc_ranlxd_uniform(&rg);
    }
    for (int axes = 0; axes < 3; axes++) {
        for (long i = 0; i < Nmesh * Nmesh * (Nmesh / 2 + 1); i++) Cdata[i] = 0;
        for (long i = 0; i < Nmesh; i++) {
            for (long j = 0; j < Nmesh; j++) {
                orc_ranlxd_seed(&rg, seedtable[i * Nmesh + j]);
                for (long k = 0; k < Nmesh / 2; k++) {
                    const double phase = orc_ranlxd_uniform(&rg) * 2 * PI;
                    double ampl;
                    do
                        ampl = orc_ranlxd_uniform(&rg);
                    while (ampl == 0);
                    if (i == Nmesh / 2 || j == Nmesh / 2 || k == Nmesh / 2) continue;
                    if (i == 0 && j == 0 && k == 0) continue;
                    double kvec[3];
                    kvec[0] = (i < Nmesh / 2) ? i * 2 * PI / Box : -(Nmesh - i) * 2 * PI / Box;
                    kvec[1] = (j < Nmesh / 2) ? j * 2 * PI / Box : -(Nmesh - j) * 2 * PI / Box;
                    kvec[2] = (k < Nmesh / 2) ? k * 2 * PI / Box : -(Nmesh - k) * 2 * PI / Box;
                    if (fabs(kvec[0]) * Box / (2 * PI) > Nsample / 2) continue;
                    if (fabs(kvec[1]) * Box / (2 * PI) > Nsample / 2) continue;
                    if (fabs(kvec[2]) * Box / (2 * PI) > Nsample / 2) continue;
                    double p_of_k = 1.0;
                    p_of_k *= -log(ampl);
                    const double delta = fac * sqrt(p_of_k) / 1.0;
                    const double re = delta * sin(phase), im = delta * cos(phase);
                    const long H = Nmesh / 2 + 1;
                    if (k > 0) {
                        Cdata[(i * Nmesh + j) * H + k] = re + I * im;
                    } else if (i == 0) {
                        if (j >= Nmesh / 2) continue;
                        const long jj = Nmesh - j;
                        Cdata[(i * Nmesh + j) * H + k] = re + I * im;
                        Cdata[(i * Nmesh + jj) * H + k] = re - I * im;
                    } else {
                        if (i >= Nmesh / 2) continue;
                        long ii = Nmesh - i, jj = Nmesh - j;
                        if (ii == Nmesh) ii = 0;
                        if (jj == Nmesh) jj = 0;
                        Cdata[(i * Nmesh + j) * H + k] = re + I * im;
                        Cdata[(ii * Nmesh + jj) * H + k] = re - I * im;
                    }
                }
            }
        }
    }
    free(seedtable);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* tabulated transfer functions                                              */
/* ------------------------------------------------------------------------ */
/* src/perturb_spline.c:30-64: for v_i = k_min + (i/acc_size)(k_max - k_min),
 * the largest j with k_j < v_i (a plain scan, as the reference) */
void orc_spline_acc_table(const double *k, int k_size, int acc_size, double *acc) {
    const double k_min = k[0], k_max = k[k_size - 1];
    for (int i = 0; i < acc_size; i++) {
        const double u = (double)i / acc_size;
        const double v = k_min + u * (k_max - k_min);
        double maxJ = 0;
        for (int j = 0; j < k_size; j++)
            if (k[j] < v) maxJ = j;
        acc[i] = maxJ;
    }
}

/* src/perturb_spline.c:73-97 */
void orc_spline_find_tau(const double *log_tau, int tau_size, double lt, int *index,
                         double *u) {
    if (lt < log_tau[0]) {
        *index = 0;
        *u = 0.0;
        return;
    }
    if (lt >= log_tau[tau_size - 1]) {
        *index = tau_size - 2;
        *u = 1.0;
        return;
    }
    for (int i = 1; i < tau_size; i++)
        if (log_tau[i] >= lt) {
            *index = i - 1;
            break;
        }
    *u = (lt - log_tau[*index]) / (log_tau[*index + 1] - log_tau[*index]);
}

/* src/perturb_spline.c:100-138 (the scan stops at the last interval, where
 * the reference would read past the table only for k outside it) */
void orc_spline_find_k(const double *k, int k_size, const double *acc, int acc_size, double kk,
                       int *index, double *u) {
    const double k_min = k[0], k_max = k[k_size - 1];
    if (kk > k_max) {
        *index = k_size - 2;
        *u = 1.0;
        return;
    }
    const double w = (kk - k_min) / (k_max - k_min);
    const int J = (int)floor(w * acc_size);
    int i = (int)acc[J < acc_size ? (J < 0 ? 0 : J) : acc_size - 1];
    for (; i < k_size - 2; i++)
        if (kk >= k[i] && kk <= k[i + 1]) break;
    *index = i;
    *u = (kk - k[i]) / (k[i + 1] - k[i]);
}

/* src/perturb_spline.c:141-160 */
double orc_spline_interp(const double *delta, int k_size, int tau_size, int k_index,
                         int tau_index, double u_k, double u_tau, int index_src) {
    const double *arr = delta + (size_t)index_src * k_size * tau_size;
    const double T11 = arr[(size_t)k_size * tau_index + k_index];
    const double T21 = arr[(size_t)k_size * tau_index + k_index + 1];
    const double T12 = arr[(size_t)k_size * (tau_index + 1) + k_index];
    const double T22 = arr[(size_t)k_size * (tau_index + 1) + k_index + 1];
    return (1 - u_tau) * ((1 - u_k) * T11 + u_k * T21) + u_tau * ((1 - u_k) * T12 + u_k * T22);
}

/* include/fft_kernels.h:131-180 inside the loop of src/fft.c:257-302 */
void orc_apply_transfer(double complex *w, const double complex *r, int N, double boxlen,
                        const double *k, int k_size, const double *delta, int tau_size,
                        int index_src, int tau_index, double u_tau, int acc_size, int invert) {
    double *acc = (double *)malloc(sizeof(double) * acc_size);
    orc_spline_acc_table(k, k_size, acc_size, acc);
    const double dk = 2 * M_PI / boxlen;
    #pragma omp parallel for
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++)
            for (int z = 0; z <= N / 2; z++) {
                double kx, ky, kz, kk, kern = 0;
                orc_wavevector(x, y, z, N, dk, &kx, &ky, &kz, &kk);
                if (kk != 0) {
                    int ki;
                    double uk;
                    orc_spline_find_k(k, k_size, acc, acc_size, kk, &ki, &uk);
                    kern = orc_spline_interp(delta, k_size, tau_size, ki, tau_index, uk, u_tau,
                                             index_src);
                    if (invert) kern = 1.0 / kern;
                }
                const size_t id = half_idx(x, y, z, N);
                const double a = creal(r[id]), b = cimag(r[id]);
                w[id] = (a * kern - b * 0.0) + I * (a * 0.0 + b * kern);
            }
    free(acc);
}

/* src/shrink_grids.c:55-75 */
int orc_shrink_grid(double *out, const double *in, int M, int N) {
    const int K = N / M;
    if (K * M != N) return 1;
    memset(out, 0, sizeof(double) * M * M * M);
    const double factor = 1.0 / (K * K * K);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
                const int x = i / K, y = j / K, z = k / K;
                out[((size_t)x * M + y) * M + z] += factor * in[((size_t)i * N + j) * N + k];
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Poisson / 2LPT                                                            */
/* ------------------------------------------------------------------------ */
/* src/poisson.c:58-70 */
int orc_solve_poisson_dg(orc_grid *g) {
    orc_fft_r2c_dg(g);
    orc_apply_kernel_dg(g, g, ORC_K_INV_POISSON, NULL);
    orc_fft_c2r_dg(g);
    return 0;
}

/* src/2lpt.c:38-143 */
int orc_solve_2lpt(orc_grid *pot, orc_grid *dens, orc_grid *work, double f1, double f2) {
    const int N = dens->N;
    const size_t chunk = (size_t)N * N * (N + 2);
    const int da[] = {ORC_K_DX, ORC_K_DY, ORC_K_DZ};
    const int ia[] = {0, 0, 0, 1, 1, 2};
    const int ib[] = {0, 1, 2, 1, 2, 2};
    if (dens->momentum_space != 1) return 1;
    orc_apply_kernel_dg(pot, dens, ORC_K_INV_POISSON, NULL);
    orc_fft_r2c_dg(dens); /* 2lpt.c:78 (redundant in the reference, kept) */
    double *hess[6];
    for (int j = 0; j < 6; j++) {
        orc_grid d;
        orc_grid_alloc(&d, N, dens->boxlen);
        orc_apply_kernel_dg(&d, pot, da[ia[j]], NULL);
        orc_apply_kernel_dg(&d, &d, da[ib[j]], NULL);
        orc_fft_c2r_dg(&d);
        free(d.fbox);
        hess[j] = d.box;
    }
    for (size_t k = 0; k < chunk; k++) {
        const double d_dxx = hess[0][k], d_dxy = hess[1][k], d_dxz = hess[2][k];
        const double d_dyy = hess[3][k], d_dyz = hess[4][k], d_dzz = hess[5][k];
        work->box[k] = d_dzz * d_dyy + d_dzz * d_dxx + d_dyy * d_dxx - d_dxy * d_dxy -
                       d_dyz * d_dyz - d_dxz * d_dxz;
    }
    for (int j = 0; j < 6; j++) free(hess[j]);
    orc_solve_poisson_dg(work);
    orc_fft_c2r_dg(pot);
    for (size_t k = 0; k < chunk; k++) pot->box[k] = pot->box[k] * f1 + f2 * work->box[k];
    orc_fft_r2c_dg(pot);
    return 0;
}

/* src/monge_ampere.c:38-42 */
static double orc_det3(const double *M) {
    return M[0] * (M[4] * M[8] - M[5] * M[7])
         - M[1] * (M[3] * M[8] - M[5] * M[6])
         + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* src/monge_ampere.c:50-183; rms_eps[c] receives the residual of cycle c */
int orc_solve_monge_ampere(orc_grid *pot, orc_grid *dens, orc_grid *work, int cycles,
                           double *rms_eps) {
    const int N = dens->N;
    const size_t chunk = (size_t)N * N * (N + 2);
    const int da[] = {ORC_K_DX, ORC_K_DY, ORC_K_DZ};
    const int ia[] = {0, 0, 0, 1, 1, 2};
    const int ib[] = {0, 1, 2, 1, 2, 2};
    if (dens->momentum_space != 1) return 1;
    orc_apply_kernel_dg(pot, dens, ORC_K_INV_POISSON, NULL);
    orc_fft_r2c_dg(dens); /* monge_ampere.c:91 (as 2lpt.c:78) */
    for (int it = 0; it < cycles; it++) {
        double *hess[6];
        for (int j = 0; j < 6; j++) {
            orc_grid d;
            orc_grid_alloc(&d, N, dens->boxlen);
            orc_apply_kernel_dg(&d, pot, da[ia[j]], NULL);
            orc_apply_kernel_dg(&d, &d, da[ib[j]], NULL);
            orc_fft_c2r_dg(&d);
            free(d.fbox);
            hess[j] = d.box;
        }
        double eps = 0, norm = 0;
        for (size_t k = 0; k < chunk; k++) {
            const double M[9] = {1 + hess[0][k], hess[1][k], hess[2][k],
                                 hess[1][k], 1 + hess[3][k], hess[4][k],
                                 hess[2][k], hess[4][k], 1 + hess[5][k]};
            const double det = orc_det3(M);
            const double rho = dens->box[k];
            const double resid = (1 + rho) - det;
            work->box[k] = resid;
            eps += resid * resid;
            norm += rho * rho;
        }
        for (int j = 0; j < 6; j++) free(hess[j]);
        orc_solve_poisson_dg(work);
        orc_fft_c2r_dg(pot);
        for (size_t k = 0; k < chunk; k++) pot->box[k] += work->box[k];
        orc_fft_r2c_dg(pot);
        if (rms_eps) rms_eps[it] = sqrt((eps / norm) / ((long long)N * N * N));
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* sigma8 normalisation of P(k) = A k^n_s T(k)^2 with a real-space top hat   */
/* ------------------------------------------------------------------------ */
static double orc_tophat_w(double x) {
    if (x < 1e-3) return 1.0 - x * x / 10.0;
    return 3.0 * (sin(x) - x * cos(x)) / (x * x * x);
}

double orc_sigma8_amplitude(double n_s, double Gamma, double sigma8, double R) {
    /* sigma^2 = int dk/(2 pi^2) k^2 P(k) W(kR)^2, Simpson's rule in ln k */
    const double lo = log(1e-6), hi = log(1e3);
    const int n = 200000;
    const double h = (hi - lo) / n;
    double sum = 0.0;
    for (int i = 0; i <= n; i++) {
        const double k = exp(lo + i * h);
        const double T = orc_bbks_transfer(k, Gamma);
        const double W = orc_tophat_w(k * R);
        const double f = k * k * k * pow(k, n_s) * T * T * W * W / (2.0 * M_PI * M_PI);
        const double w = (i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
        sum += w * f;
    }
    const double I1 = sum * h / 3.0;
    return sigma8 * sigma8 / I1;
}

/* ------------------------------------------------------------------------ */
/* particles                                                                 */
/* ------------------------------------------------------------------------ */
/* src/particle.c:87-129 with FirstID from src/particle_types.c:102-104 */
void orc_lattice(int M, double boxlen, long long first_id, float *X, float *Y, float *Z,
                 long long *id) {
    float len = (float)boxlen;
    float spacing = len / M;
    long long c = 0;
    for (int x = 0; x < M; x++)
        for (int y = 0; y < M; y++)
            for (int z = 0; z < M; z++) {
                if (id) id[c] = ((long long)x * M * M + (long long)y * M + z) + first_id;
                X[c] = ORC_WRAP(x, M) * spacing;
                Y[c] = ORC_WRAP(y, M) * spacing;
                Z[c] = ORC_WRAP(z, M) * spacing;
                c++;
            }
}

/* src/grids_interp.c:180-207 on one rank: padded slice, every index wrapped */
static inline double orc_access(const double *box, int iX, int iY, int iZ, int N) {
    iX = ORC_WRAP(iX, N);
    iY = ORC_WRAP(iY, N);
    iZ = ORC_WRAP(iZ, N);
    return box[((size_t)iX * N + iY) * (N + 2) + iZ];
}

/* src/grids_interp.c:267-309 */
double orc_grid_tsc(const double *box, int N, double boxlen, double x, double y, double z) {
    double X = x * N / boxlen;
    double Y = y * N / boxlen;
    double Z = z * N / boxlen;
    int iX = (int)floor(X);
    int iY = (int)floor(Y);
    int iZ = (int)floor(Z);
    double lookLength = 1.5;
    int lookLftX = (int)floor((X - iX) - lookLength);
    int lookRgtX = (int)floor((X - iX) + lookLength);
    int lookLftY = (int)floor((Y - iY) - lookLength);
    int lookRgtY = (int)floor((Y - iY) + lookLength);
    int lookLftZ = (int)floor((Z - iZ) - lookLength);
    int lookRgtZ = (int)floor((Z - iZ) + lookLength);
    double sum = 0;
    for (int i = lookLftX; i <= lookRgtX; i++) {
        for (int j = lookLftY; j <= lookRgtY; j++) {
            for (int k = lookLftZ; k <= lookRgtZ; k++) {
                double xx = fabs(X - (iX + i));
                double yy = fabs(Y - (iY + j));
                double zz = fabs(Z - (iZ + k));
                double part_x = xx < 0.5 ? (0.75 - xx * xx) : (xx < 1.5 ? 0.5 * (1.5 - xx) * (1.5 - xx) : 0);
                double part_y = yy < 0.5 ? (0.75 - yy * yy) : (yy < 1.5 ? 0.5 * (1.5 - yy) * (1.5 - yy) : 0);
                double part_z = zz < 0.5 ? (0.75 - zz * zz) : (zz < 1.5 ? 0.5 * (1.5 - zz) * (1.5 - zz) : 0);
                sum += orc_access(box, iX + i, iY + j, iZ + k, N) * (part_x * part_y * part_z);
            }
        }
    }
    return sum;
}

/* src/mitos.c:827-863 (components in sequence, each at the updated position)
 * then the periodic wrap of :1072-1077 */
void orc_displace_tsc(int N, double boxlen, const double *psi_x, const double *psi_y,
                      const double *psi_z, long long npart, float *X, float *Y, float *Z) {
    const double *psi[3] = {psi_x, psi_y, psi_z};
    for (int dir = 0; dir < 3; dir++) {
        #pragma omp parallel for
        for (long long i = 0; i < npart; i++) {
            double x = X[i], y = Y[i], z = Z[i];
            double disp = orc_grid_tsc(psi[dir], N, boxlen, x, y, z);
            if (dir == 0) X[i] -= disp;
            else if (dir == 1) Y[i] -= disp;
            else Z[i] -= disp;
        }
    }
    #pragma omp parallel for
    for (long long i = 0; i < npart; i++) {
        X[i] = orc_fwrap(X[i], boxlen);
        Y[i] = orc_fwrap(Y[i], boxlen);
        Z[i] = orc_fwrap(Z[i], boxlen);
    }
}

/* src/mitos.c:827-863 (moves), :867-907 (velocities at the moved positions,
 * before the wrap), :1072-1077 (wrap); psi or vel may be NULL */
void orc_displace_velocity_tsc(int N, double boxlen, const double *const *psi,
                               const double *const *vel, long long npart, float *X, float *Y,
                               float *Z, float *VX, float *VY, float *VZ) {
    if (psi) {
        for (int dir = 0; dir < 3; dir++) {
            #pragma omp parallel for
            for (long long i = 0; i < npart; i++) {
                double x = X[i], y = Y[i], z = Z[i];
                double disp = orc_grid_tsc(psi[dir], N, boxlen, x, y, z);
                if (dir == 0) X[i] -= disp;
                else if (dir == 1) Y[i] -= disp;
                else Z[i] -= disp;
            }
        }
    }
    if (vel) {
        for (int dir = 0; dir < 3; dir++) {
            #pragma omp parallel for
            for (long long i = 0; i < npart; i++) {
                double x = X[i], y = Y[i], z = Z[i];
                double v = orc_grid_tsc(vel[dir], N, boxlen, x, y, z);
                if (dir == 0) VX[i] = v;
                else if (dir == 1) VY[i] = v;
                else VZ[i] = v;
            }
        }
    }
    #pragma omp parallel for
    for (long long i = 0; i < npart; i++) {
        X[i] = orc_fwrap(X[i], boxlen);
        Y[i] = orc_fwrap(Y[i], boxlen);
        Z[i] = orc_fwrap(Z[i], boxlen);
    }
}

/* analyse/src/mitos_read.c:278-334 */
double orc_cic(int N, double boxlen, long long npart, const float *Xp, const float *Yp,
               const float *Zp, double mass, double *box) {
    const double grid_cell_vol = boxlen * boxlen * boxlen / ((double)N * N * N);
    double total_mass = 0;
    for (long long l = 0; l < npart; l++) {
        double x = orc_fwrap(Xp[l], boxlen);
        double y = orc_fwrap(Yp[l], boxlen);
        double z = orc_fwrap(Zp[l], boxlen);
        double X = x / (boxlen / N);
        double Y = y / (boxlen / N);
        double Z = z / (boxlen / N);
        double M = mass;
        total_mass += M;
        int iX = (int)floor(X);
        int iY = (int)floor(Y);
        int iZ = (int)floor(Z);
        double dx = X - iX, dy = Y - iY, dz = Z - iZ;
        double tx = 1.0 - dx, ty = 1.0 - dy, tz = 1.0 - dz;
        double val = M / grid_cell_vol;
#define RM(a, b, c) (((size_t)ORC_WRAP(a, N) * N + ORC_WRAP(b, N)) * N + ORC_WRAP(c, N))
        box[RM(iX, iY, iZ)] += val * tx * ty * tz;
        box[RM(iX + 1, iY, iZ)] += val * dx * ty * tz;
        box[RM(iX, iY + 1, iZ)] += val * tx * dy * tz;
        box[RM(iX, iY, iZ + 1)] += val * tx * ty * dz;
        box[RM(iX + 1, iY + 1, iZ)] += val * dx * dy * tz;
        box[RM(iX + 1, iY, iZ + 1)] += val * dx * ty * dz;
        box[RM(iX, iY + 1, iZ + 1)] += val * tx * dy * dz;
        box[RM(iX + 1, iY + 1, iZ + 1)] += val * dx * dy * dz;
#undef RM
    }
    return total_mass;
}

/* analyse/src/mitos_read.c:363-376 */
void orc_overdensity(int N, double boxlen, double total_mass, double *box) {
    const double avg = total_mass / (boxlen * boxlen * boxlen);
    const size_t tot = (size_t)N * N * N;
    for (size_t i = 0; i < tot; i++) box[i] = (box[i] - avg) / avg;
}

/* src/calc_powerspec.c:32-90 */
void orc_powerspec(int N, double boxlen, const double complex *box1, const double complex *box2,
                   int bins, double *k_in_bins, double *power_in_bins, int *obs_in_bins) {
    const double boxvol = boxlen * boxlen * boxlen;
    const double dk = 2 * M_PI / boxlen;
    const double max_k = sqrt(3) * dk * N / 2;
    const double min_k = dk;
    const double log_max_k = log(max_k);
    const double log_min_k = log(min_k);
    for (int i = 0; i < bins; i++) {
        k_in_bins[i] = 0;
        power_in_bins[i] = 0;
        obs_in_bins[i] = 0;
    }
    double kx, ky, kz, k;
    for (int x = 0; x < N; x++) {
        for (int y = 0; y < N; y++) {
            for (int z = 0; z <= N / 2; z++) {
                orc_wavevector(x, y, z, N, dk, &kx, &ky, &kz, &k);
                if (k == 0) continue;
                const float u = (log(k) - log_min_k) / (log_max_k - log_min_k);
                const int bin = floor((bins - 1) * u);
                const size_t id = half_idx(x, y, z, N);
                if (bin < 0 || bin >= bins) continue; /* reference asserts */
                double a1 = creal(box1[id]), a2 = creal(box2[id]);
                double b1 = cimag(box1[id]), b2 = cimag(box2[id]);
                double Power = a1 * a2 + b1 * b2;
                int multiplicity = (z == 0 || z == N / 2) ? 1 : 2;
                k_in_bins[bin] += multiplicity * k;
                power_in_bins[bin] += multiplicity * Power;
                obs_in_bins[bin] += multiplicity;
            }
        }
    }
    for (int i = 0; i < bins; i++) {
        k_in_bins[i] /= obs_in_bins[i];
        power_in_bins[i] /= obs_in_bins[i];
        power_in_bins[i] /= boxvol;
    }
}

/* ------------------------------------------------------------------------ */
/* full pipeline                                                             */
/* ------------------------------------------------------------------------ */
#include <time.h>
static double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int orc_ic_run(const orc_ic_config *cfg, orc_ic_outputs *out) {
    const int N = cfg->N;
    const double L = cfg->boxlen;
    const size_t half = (size_t)N * N * (N / 2 + 1);
    const size_t padded = (size_t)N * N * (N + 2);
    const long long np = (long long)N * N * N;
    const int M = cfg->M > 0 ? cfg->M : N; /* particle lattice, src/particle.c:87-129 */
    const long long npart = (long long)M * M * M;
    double t0 = orc_now(), t1;
    for (int i = 0; i < 8; i++) out->stage_seconds[i] = 0;

    /* amplitude of P(k) = A k^n_s T_BBKS(k)^2 normalised to sigma8 (R = 8 Mpc/h) */
    const double A = orc_sigma8_amplitude(cfg->n_s, cfg->gamma, cfg->sigma8, 8.0);

    /* src/mitos.c:148 (rank 0) */
    orc_rng rng = orc_rng_init(cfg->seed);

    /* src/mitos.c:216-227: GRF, Hermitian symmetry, primordial kernel, c2r */
    orc_grid grf;
    if (orc_grid_alloc(&grf, N, L)) return 10;
    orc_generate_complex_grf(&grf, &rng);
    orc_enforce_hermiticity(&grf);
    const double pl[2] = {A, cfg->n_s};
    orc_apply_kernel_dg(&grf, &grf, ORC_K_SQRT_POWER_LAW, pl);
    orc_fft_c2r_dg(&grf);
    orc_fft_r2c_dg(&grf); /* :300 */
    t1 = orc_now(); out->stage_seconds[0] = t1 - t0; t0 = t1;

    orc_grid grid, pot, der;
    if (orc_grid_alloc(&grid, N, L) || orc_grid_alloc(&pot, N, L) || orc_grid_alloc(&der, N, L))
        return 10;

    /* src/generate_grids.c:52-60: transfer kernel then c2r; src/mitos.c:377 r2c */
    const double gp[1] = {cfg->gamma};
    orc_apply_kernel_dg(&grid, &grf, ORC_K_BBKS, gp);
    orc_fft_c2r_dg(&grid);
    orc_fft_r2c_dg(&grid);
    if (out->delta_k) memcpy(out->delta_k, grid.fbox, sizeof(double complex) * half);
    orc_grid_free(&grf);
    t1 = orc_now(); out->stage_seconds[1] = t1 - t0; t0 = t1;

    /* src/mitos.c:380-391 */
    if (cfg->run2lpt) {
        int err = orc_solve_2lpt(&pot, &grid, &der, cfg->f1, cfg->f2);
        if (err) return err;
    } else {
        orc_apply_kernel_dg(&pot, &grid, ORC_K_INV_POISSON, NULL);
    }
    t1 = orc_now(); out->stage_seconds[2] = t1 - t0; t0 = t1;

    /* src/mitos.c:394-416: undo TSC window, three derivative grids */
    const double hw[3] = {3, N, L};
    orc_apply_kernel_dg(&pot, &pot, ORC_K_UNDO_WINDOW, hw);
    double *psi = out->psi ? out->psi : (double *)malloc(sizeof(double) * 3 * padded);
    const int dker[3] = {ORC_K_DX, ORC_K_DY, ORC_K_DZ};
    for (int i = 0; i < 3; i++) {
        orc_apply_kernel_dg(&der, &pot, dker[i], NULL);
        orc_fft_c2r_dg(&der);
        memcpy(psi + i * padded, der.box, sizeof(double) * padded);
    }
    orc_grid_free(&grid);
    orc_grid_free(&pot);
    orc_grid_free(&der);
    t1 = orc_now(); out->stage_seconds[3] = t1 - t0; t0 = t1;

    /* src/particle.c:87-129; src/mitos.c:827-863, :1072-1077 */
    float *X = out->X ? out->X : (float *)malloc(sizeof(float) * npart);
    float *Y = out->Y ? out->Y : (float *)malloc(sizeof(float) * npart);
    float *Z = out->Z ? out->Z : (float *)malloc(sizeof(float) * npart);
    orc_lattice(M, L, 1, X, Y, Z, out->id);
    orc_displace_tsc(N, L, psi, psi + padded, psi + 2 * padded, npart, X, Y, Z);
    if (!out->psi) free(psi);
    t1 = orc_now(); out->stage_seconds[4] = t1 - t0; t0 = t1;

    /* analyse/src/mitos_read.c:278-376 */
    double *dens = out->density ? out->density : (double *)calloc(np, sizeof(double));
    if (out->density) memset(dens, 0, sizeof(double) * np);
    double total = orc_cic(N, L, npart, X, Y, Z, 1.0, dens);
    orc_overdensity(N, L, total, dens);
    if (!out->X) free(X);
    if (!out->Y) free(Y);
    if (!out->Z) free(Z);
    t1 = orc_now(); out->stage_seconds[5] = t1 - t0; t0 = t1;

    /* analyse/src/mitos_read.c:409-419 */
    double complex *fbox = (double complex *)malloc(sizeof(double complex) * half);
    orc_fft_r2c(N, dens, N, fbox);
    orc_normalize_r2c(fbox, N, L);
    const double cw[3] = {2, N, L};
    orc_apply_kernel(fbox, fbox, N, L, ORC_K_UNDO_WINDOW, cw);
    if (out->k_in_bins)
        orc_powerspec(N, L, fbox, fbox, cfg->bins, out->k_in_bins, out->power_in_bins,
                      out->obs_in_bins);
    free(fbox);
    if (!out->density) free(dens);
    t1 = orc_now(); out->stage_seconds[6] = t1 - t0;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* helpers for the checks                                                    */
/* ------------------------------------------------------------------------ */
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* per-mode reference bin of calc_cross_powerspec for (x,y,z) -- for LUT checks */
int orc_pk_bin(int N, double boxlen, int bins, int x, int y, int z) {
    const double dk = 2 * M_PI / boxlen;
    const double max_k = sqrt(3) * dk * N / 2;
    const double min_k = dk;
    const double log_max_k = log(max_k);
    const double log_min_k = log(min_k);
    double kx, ky, kz, k;
    orc_wavevector(x, y, z, N, dk, &kx, &ky, &kz, &k);
    if (k == 0) return -1;
    const float u = (log(k) - log_min_k) / (log_max_k - log_min_k);
    return (int)floor((bins - 1) * u);
}

/* number of half-grid modes whose reference bin differs from lut[n^2] */
long long orc_pk_lut_mismatches(int N, double boxlen, int bins, const int *lut) {
    long long bad = 0;
    #pragma omp parallel for reduction(+ : bad)
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++)
            for (int z = 0; z <= N / 2; z++) {
                const int nx = x > N / 2 ? x - N : x;
                const int ny = y > N / 2 ? y - N : y;
                const long n2 = (long)nx * nx + (long)ny * ny + (long)z * z;
                if (n2 == 0) continue;
                if (orc_pk_bin(N, boxlen, bins, x, y, z) != lut[n2]) bad++;
            }
    return bad;
}
