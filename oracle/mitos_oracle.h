/*
 * CPU oracle for the Mitos initial-condition hot path (FP64, plain C99).
 *
 * TEST INFRASTRUCTURE ONLY.  This file and mitos_oracle.c restate the
 * reference algorithm so that tests, the smoke check and the CPU-baseline leg
 * of the benchmark can check the HIP library against it.  Nothing in the
 * product path (libmitos_hip.so, the mitos_amd package) may link, load or call
 * it.
 *
 * Each function cites the reference file:line it follows.  FFTW is not part
 * of the reference tree; its r2c/c2r conventions (unnormalised DFT, e^{-2pi i}
 * forward, last axis half-complex, imaginary parts of the last-axis DC and
 * Nyquist entries ignored by c2r) are restated by the oracle's own FFT.
 */
#ifndef MITOS_ORACLE_H
#define MITOS_ORACLE_H

#include <stdint.h>
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: include/random_xorshift.h:7-59, include/random.h:25-33 ---- */
typedef struct { uint64_t s[4]; } orc_rng;

orc_rng orc_rng_init(uint64_t seed);
uint64_t orc_rng_next(orc_rng *st);
/* n raw draws after skipping `skip` draws from the seed's start */
void orc_rng_draws(uint64_t seed, uint64_t skip, uint64_t n, uint64_t *out);
void orc_rng_draws_at(uint64_t seed, const uint64_t *offs, uint64_t n, uint64_t k,
                      uint64_t *out);
void orc_rng_state_after(uint64_t seed, uint64_t n, uint64_t out[4]);
/* src/random.c:34-49 */
double orc_sample_norm(orc_rng *st);
/* n Box-Muller normals starting at the current state (advances it) */
void orc_sample_norms(uint64_t seed, uint64_t skip_draws, uint64_t n, double *out);

/* ---- grid: include/distributed_grid.h:29-60 (one rank: NX=N, X0=0) ---- */
typedef struct {
    int N;
    double boxlen;
    char momentum_space;
    double complex *fbox; /* N*N*(N/2+1) */
    double *box;          /* N*N*(N+2), padded real layout */
} orc_grid;

int orc_grid_alloc(orc_grid *g, int N, double boxlen);
void orc_grid_free(orc_grid *g);

/* src/fft.c:29-35 */
void orc_wavevector(int x, int y, int z, int N, double dk,
                    double *kx, double *ky, double *kz, double *k);

/* ---- GRF: src/grf.c:26-69, 78-175 ---- */
int orc_generate_complex_grf(orc_grid *g, orc_rng *st);
int orc_enforce_hermiticity(orc_grid *g);

/* ---- FFT (FFTW semantics), unnormalised ----
 * r2c: real (row stride rs >= N) -> half complex N*N*(N/2+1)
 * c2r: half complex (not modified) -> real (row stride rs); padding untouched */
void orc_fft_r2c(int N, const double *in, int rs, double complex *out);
void orc_fft_c2r(int N, const double complex *in, double *out, int rs);

/* src/fft.c:219-254 (normalised, on the padded grid) */
int orc_fft_r2c_dg(orc_grid *g);
int orc_fft_c2r_dg(orc_grid *g);

/* ---- kernels: include/fft_kernels.h (by id) ---- */
enum {
    ORC_K_CONSTANT = 0,        /* params[0] = constant            (:28-31)  */
    ORC_K_INV_POISSON = 1,     /*                                 (:96-99)  */
    ORC_K_DX = 2, ORC_K_DY = 3, ORC_K_DZ = 4, /*                  (:101-114)*/
    ORC_K_POWER_NO_TRANSFER = 5, /* params A_s, n_s, k_pivot      (:116-121, src/primordial.c:25-33) */
    ORC_K_UNDO_WINDOW = 6,     /* params order, N, boxlen         (:198-219)*/
    ORC_K_SQRT_POWER_LAW = 7,  /* params A, n_s: sqrt(A k^n_s), 0 at k=0 */
    ORC_K_BBKS = 8,            /* params Gamma: T_BBKS(k), 0 at k=0 */
    ORC_K_LOWPASS = 9,         /* params k_max                    (:33-37)  */
    ORC_K_HIPASS = 10,         /* params k_min                    (:39-43)  */
    ORC_K_TOPHAT = 11,         /* params k_min, k_max             (:45-51)  */
    ORC_K_ELLIPTIC_TOPHAT = 12, /* params rx, ry, rz, eta2        (:54-68)  */
    ORC_K_TRANSLATE = 13,      /* params rx, ry, rz               (:70-79)  */
    ORC_K_GAUSSIAN = 14,       /* params R                        (:81-86)  */
    ORC_K_GAUSSIAN_INV = 15    /* params R                        (:88-93)  */
};
double orc_kernel_value_re(int kid, const double *params, double kx, double ky,
                           double kz, double k, double *im);
/* src/fft.c:257-302 */
int orc_apply_kernel_dg(orc_grid *w, const orc_grid *r, int kid, const double *params);
/* src/fft.c:134-159 (serial unpadded arrays) */
int orc_apply_kernel(double complex *w, const double complex *r, int N, double boxlen,
                     int kid, const double *params);
int orc_normalize_r2c(double complex *arr, int N, double boxlen); /* src/fft.c:38-49 */

/* ---- RANLUX ranlxd1 and the N-GenIC noise: src/grf_ngeniclike.c:28-203 ---- */
typedef struct {
    double x[12];
    double carry;
    int ir, jr, ir_old, pr;
} orc_ranlxd;
void orc_ranlxd_draws(unsigned long seed, long n, double *out);
int orc_generate_ngeniclike_grf(double complex *fbox, int N, double boxlen, long seed);

/* ---- tabulated transfer functions: src/perturb_spline.c:30-169 ---- */
/* acceleration table of initPerturbSpline (:30-64) */
void orc_spline_acc_table(const double *k, int k_size, int acc_size, double *acc);
/* perturbSplineFindTau (:73-97), perturbSplineFindK (:100-138) */
void orc_spline_find_tau(const double *log_tau, int tau_size, double lt, int *index,
                         double *u);
void orc_spline_find_k(const double *k, int k_size, const double *acc, int acc_size, double kk,
                       int *index, double *u);
/* perturbSplineInterp (:141-160) */
double orc_spline_interp(const double *delta, int k_size, int tau_size, int k_index,
                         int tau_index, double u_k, double u_tau, int index_src);
/* kernel_(inv_)transfer_function (include/fft_kernels.h:131-180) applied to
 * the half grid r (src/fft.c:257-302) */
void orc_apply_transfer(double complex *w, const double complex *r, int N, double boxlen,
                        const double *k, int k_size, const double *delta, int tau_size,
                        int index_src, int tau_index, double u_tau, int acc_size, int invert);
/* shrinkGrid (src/shrink_grids.c:55-75), unpadded N^3 -> M^3 */
int orc_shrink_grid(double *out, const double *in, int M, int N);

/* ---- Poisson / 2LPT: src/poisson.c:58-70, src/2lpt.c:38-143 ---- */
int orc_solve_poisson_dg(orc_grid *g);
int orc_solve_2lpt(orc_grid *pot, orc_grid *dens, orc_grid *work, double f1, double f2);
/* src/monge_ampere.c:50-183; rms_eps (may be NULL) gets one value per cycle */
int orc_solve_monge_ampere(orc_grid *pot, orc_grid *dens, orc_grid *work, int cycles,
                           double *rms_eps);

/* ---- spectrum: BBKS transfer function, sigma8 normalisation ---- */
double orc_bbks_transfer(double k, double Gamma);
double orc_sigma8_amplitude(double n_s, double Gamma, double sigma8, double R);

/* ---- particles ---- */
/* src/particle.c:87-129, src/particle_types.c:102-104 */
void orc_lattice(int M, double boxlen, long long first_id,
                 float *X, float *Y, float *Z, long long *id);
/* src/grids_interp.c:267-309 with access_grid :180-207 on one rank */
double orc_grid_tsc(const double *box_padded, int N, double boxlen,
                    double x, double y, double z);
/* src/mitos.c:827-863 then :1072-1077 */
void orc_displace_velocity_tsc(int N, double boxlen, const double *const *psi,
                               const double *const *vel, long long npart, float *X, float *Y,
                               float *Z, float *VX, float *VY, float *VZ);
void orc_displace_tsc(int N, double boxlen, const double *psi_x, const double *psi_y,
                      const double *psi_z, long long npart, float *X, float *Y, float *Z);
/* analyse/src/mitos_read.c:278-334 (double CIC into unpadded N^3 grid);
 * returns the total mass */
double orc_cic(int N, double boxlen, long long npart, const float *X, const float *Y,
               const float *Z, double mass, double *box);
/* analyse/src/mitos_read.c:363-376 */
void orc_overdensity(int N, double boxlen, double total_mass, double *box);
/* src/calc_powerspec.c:32-90 */
void orc_powerspec(int N, double boxlen, const double complex *b1, const double complex *b2,
                   int bins, double *k_in_bins, double *power_in_bins, int *obs_in_bins);

/* ---- full pipeline in the reference's call order ----
 * src/mitos.c:148,205-228,300,352-423,768-863,1072-1077 and
 * analyse/src/mitos_read.c:278-434.  Outputs are optional (NULL = skip). */
typedef struct {
    int N;             /* grid size */
    double boxlen;
    uint64_t seed;
    double n_s;
    double sigma8;
    double gamma;      /* BBKS shape parameter */
    double f1, f2;     /* 2LPT factors (1, -3/7) */
    int run2lpt;
    int bins;
    int M;             /* particle lattice M^3 (src/particle.c:87-129); 0: M = N */
} orc_ic_config;

typedef struct {
    double *psi;       /* 3 * N*N*(N+2) displacement grids x,y,z (padded) */
    float *X, *Y, *Z;  /* M^3 particle positions after wrap */
    long long *id;     /* M^3 particle ids */
    double *density;   /* N^3 overdensity (unpadded) */
    double *k_in_bins, *power_in_bins;
    int *obs_in_bins;
    double complex *delta_k; /* N*N*(N/2+1) normalised density field before 2LPT */
    double stage_seconds[8];
} orc_ic_outputs;

int orc_ic_run(const orc_ic_config *cfg, orc_ic_outputs *out);
/* OpenMP threads of the loops (orc_set_threads) and of the FFTs alone
 * (orc_set_fft_threads: 0 = the same as the loops, 1 = serial as FFTW's
 * plans in the reference) */
int orc_set_threads(int n);
int orc_set_fft_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
