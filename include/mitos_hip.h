/*
 * mitos_hip.h -- C ABI of libmitos_hip.so, an MI355X (gfx950) implementation
 * of the Mitos initial-condition hot path.
 *
 * Part 1 keeps the reference's own call surface (names, argument meaning,
 * return codes) so a reference caller relinks against this library on one
 * GPU.  Every reference-named function cites the declaration it replaces.
 * Part 2 adds the batched particle entry points that replace the reference's
 * per-particle loops, and Part 3 the fused whole-pipeline, ensemble and
 * adjoint entry points.
 *
 * Memory: grid arrays (dg->fbox, dg->box and every array argument documented
 * as "device") live in GPU memory.  Host code reads or writes them through
 * mitos_grid_to_host / mitos_grid_from_host.  All Part 1 and Part 2 calls are
 * host-synchronous; they return 0 on success and a MITOS_ERR_* code
 * otherwise (the library never calls exit()).
 *
 * One rank: NX = N, X0 = 0, local_size = N*N*(N/2+1).  Grid sizes with compiled
 * FFT passes: 8, 16, 32, 64, 128, 256, 512, 1024 and 24, 48, 96, 192, 384, 768,
 * 1536.
 */
#ifndef MITOS_HIP_H
#define MITOS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#pragma GCC visibility push(default)

/* ---- types --------------------------------------------------------------- */

/* FFTW's complex types (include/fft.h:23-24 includes complex.h before fftw3.h,
 * so the reference sees double _Complex); identical layout to double[2]. */
#if !defined(FFTW3_H)
#ifdef __cplusplus
typedef _Complex double fftw_complex;
typedef _Complex float fftwf_complex;
#else
typedef double _Complex fftw_complex;
typedef float _Complex fftwf_complex;
#endif
#endif

/* MPI communicator: only stored (one rank).  MPICH's integer handle unless
 * mpi.h was included first. */
#if !defined(MPI_VERSION)
typedef int MPI_Comm;
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#endif
#endif

/* error codes: 1 and 2 as fft_apply_kernel_dg returns them (src/fft.c:269-277);
 * solve2LPT returns 1 for a density not in momentum space (src/2lpt.c:69-72) */
#define MITOS_OK 0
#define MITOS_ERR_DIMS 1
#define MITOS_ERR_NOT_MOMENTUM 2
#define MITOS_ERR_DEVICE 3
#define MITOS_ERR_UNSUPPORTED 4
#define MITOS_ERR_ARG 5
#define MITOS_ERR_ALLOC 6
#define MITOS_ERR_REAL_SPACE 7
#define MITOS_ERR_TOO_LARGE 8 /* input beyond a fixed scratch of an allocation-free call */

/* include/distributed_grid.h:29-60 -- same fields, same order */
struct distributed_grid {
    int N;
    double boxlen;
    MPI_Comm comm;
    char momentum_space;
    long int NX;
    long int X0;
    long int local_size;
    fftw_complex *fbox; /* device: N*N*(N/2+1) */
    double *box;        /* device: N*N*(N+2), padded along z */
};

/* include/fft.h:31-39 */
struct kernel {
    double kx, ky, kz;
    double k;
#ifdef __cplusplus
    _Complex double kern;
#else
    double _Complex kern;
#endif
    const void *params;
};

/* include/input.h:124-137 (fields read by kernel_power_no_transfer) */
struct cosmology {
    double h;
    double n_s;
    double A_s;
    double k_pivot;
    double z_ini;
    double log_tau_ini;
    double rho_crit;
    double z_source;
    double log_tau_source;
};

/* include/fft_kernels.h:189-193 */
struct Hermite_kern_params {
    int order;
    int N;
    double boxlen;
};

/* include/random_xorshift.h:11-13, include/random.h:25 */
struct xoshiro256ss_state {
    uint64_t s[4];
};
typedef struct xoshiro256ss_state rng_state;

/* ======================================================================== */
/* Part 1: reference call surface                                           */
/* ======================================================================== */

/* include/distributed_grid.h:74-77 -- device allocation of fbox and box */
int alloc_local_grid(struct distributed_grid *dg, int N, double boxlen, MPI_Comm comm);
int free_local_grid(struct distributed_grid *dg);
int free_local_real_grid(struct distributed_grid *dg);
int free_local_complex_grid(struct distributed_grid *dg);

/* host access to a grid's device arrays (either pointer may be NULL) */
int mitos_grid_to_host(const struct distributed_grid *dg, fftw_complex *fbox_host,
                       double *box_host);
int mitos_grid_from_host(struct distributed_grid *dg, const fftw_complex *fbox_host,
                         const double *box_host);

/* include/random.h:27-33 (host) */
uint64_t rand_uint64(rng_state *state);
rng_state rand_uint64_init(uint64_t seed);
/* include/random.h:91-94 (host; Box-Muller with the host libm, cosine
 * branch, two draws per normal, as src/random.c:34-49) */
double sampleUniform(rng_state *state);
double sampleNorm(rng_state *state);

/* include/grf.h:30-31 -- the state advances by the 4*(N*N*(N/2+1)-1) draws
 * the reference consumes, so later draws continue the same stream */
int generate_complex_grf(struct distributed_grid *dg, rng_state *state);
int enforce_hermiticity(struct distributed_grid *dg);
/* include/grf_ngeniclike.h:29-30 -- the N-GenIC white noise of
 * src/grf_ngeniclike.c:28-203 into the device slab [X0, X0+NX) of an
 * N x N x (N/2+1) half grid (block_size unused, as in the reference): a
 * RANLUX (ranlxd1) stream seeded with `seed` fills the column seed table,
 * each column reseeds and draws its phases and amplitudes; Nyquist modes,
 * DC and the z = N/2 plane are zero.  Returns 4 when the seed table cannot
 * be allocated (src/grf_ngeniclike.c:68-69). */
int generate_ngeniclike_grf(fftw_complex *fbox, int N, int NX, int X0, long int block_size,
                            double boxlen, long int seed);

/* include/fft.h:90-116 */
void fft_wavevector(int x, int y, int z, int N, double delta_k, double *kx, double *ky,
                    double *kz, double *k);
int fft_normalize_r2c(fftw_complex *arr, int N, double boxlen);             /* device arr */
int fft_normalize_r2c_float(fftwf_complex *arr, int N, double boxlen);      /* device arr */
int fft_normalize_r2c_float_mpi(fftwf_complex *farr, long int N, long int local_size,
                                double boxlen);                             /* device arr */
int fft_normalize_c2r(double *arr, int N, double boxlen);                   /* device N^3 */
int fft_apply_kernel(fftw_complex *write, const fftw_complex *read, int N, double boxlen,
                     void (*compute)(struct kernel *the_kernel), const void *params);
int fft_apply_kernel_float(fftwf_complex *write, const fftwf_complex *read, int N,
                           double boxlen, void (*compute)(struct kernel *the_kernel),
                           const void *params);
int fft_apply_kernel_float_mpi(fftwf_complex *write, const fftwf_complex *read, long int N,
                               long int X0, long int NX, double boxlen,
                               void (*compute)(struct kernel *the_kernel),
                               const void *params);
int fft_r2c_dg(struct distributed_grid *dg);
int fft_c2r_dg(struct distributed_grid *dg);
int fft_apply_kernel_dg(struct distributed_grid *dg_write, const struct distributed_grid *dg_read,
                        void (*compute)(struct kernel *the_kernel), const void *params);

/* Kernel callbacks (include/fft_kernels.h:28-180, 198-219).  They
 * evaluate the kernel on the host when called directly; passed to an
 * fft_apply_kernel* function, their address selects the device kernel.  Any
 * other callback -- a caller's own kernel, or the static inline copy of a
 * reference kernel that include/fft_kernels.h puts in every translation
 * unit -- is evaluated on the host for every mode (threads over x-slabs, as
 * the OpenMP loop of src/fft.c:279-291), and the device multiplies the grid
 * by those values.  That path is host-bound; the replacement
 * include/fft_kernels.h of this library declares the kernels extern, so a
 * unit compiled against it reaches the device kernels. */
void kernel_constant(struct kernel *the_kernel);          /* params: const double* */
void kernel_lowpass(struct kernel *the_kernel);           /* params: double k_max  */
void kernel_hipass(struct kernel *the_kernel);            /* params: double k_min  */
void kernel_tophat(struct kernel *the_kernel);            /* params: double[2]     */
void kernel_elliptic_tophat(struct kernel *the_kernel);   /* params: double[4]     */
void kernel_translate(struct kernel *the_kernel);         /* params: double[3]     */
void kernel_gaussian(struct kernel *the_kernel);          /* params: double R      */
void kernel_gaussian_inv(struct kernel *the_kernel);      /* params: double R      */
void kernel_inv_poisson(struct kernel *the_kernel);       /* params: unused        */
void kernel_dx(struct kernel *the_kernel);
void kernel_dy(struct kernel *the_kernel);
void kernel_dz(struct kernel *the_kernel);
void kernel_power_no_transfer(struct kernel *the_kernel); /* params: struct cosmology* */
void kernel_undo_Hermite_window(struct kernel *the_kernel); /* params: Hermite_kern_params* */
/* closed-form spectra (new): params point to struct mitos_spectrum_params */
struct mitos_spectrum_params {
    double A;     /* amplitude of P(k) = A k^n_s T(k)^2 */
    double n_s;
    double Gamma; /* BBKS shape parameter (k in h/Mpc) */
};
void kernel_sqrt_power_law(struct kernel *the_kernel); /* sqrt(A k^n_s), 0 at k=0 */
void kernel_bbks_transfer(struct kernel *the_kernel);  /* T_BBKS(k/Gamma), 0 at k=0 */

/* ---- tabulated transfer functions T(k, log tau) ------------------------- */
/* include/perturb_data.h:25-52 -- same fields, same order (host arrays) */
struct perturb_data {
    int k_size;
    int tau_size;
    int n_functions;
    double *delta;     /* k_size * tau_size * n_functions, k fastest */
    double *k;         /* k_size, increasing */
    double *log_tau;   /* tau_size, increasing */
    double *redshift;
    double *D_growth;
    double *f_growth;
    double *H_Hubble;
    char **titles;     /* n_functions */
    double *Omega;
};
/* include/perturb_spline.h:25-34 */
struct perturb_spline {
    const struct perturb_data *ptdat;
    double *k_acc_table;
    int k_acc_table_size;
};
/* include/fft_kernels.h:123-128: params of kernel_(inv_)transfer_function */
struct spline_params {
    const struct perturb_spline *spline;
    int index_src;
    int tau_index;
    double u_tau;
};

/* src/perturb_spline.c:30-169 (host) */
int initPerturbSpline(struct perturb_spline *spline, int k_acc_size,
                      const struct perturb_data *ptdat);
int cleanPerturbSpline(struct perturb_spline *spline);
int perturbSplineFindTau(const struct perturb_spline *spline, double log_tau, int *index,
                         double *u);
int perturbSplineFindK(const struct perturb_spline *spline, double k, int *index, double *u);
double perturbSplineInterp(const struct perturb_spline *spline, int k_index, int tau_index,
                           double u_k, double u_tau, int index_src);
double perturbSplineInterp0(const struct perturb_spline *spline, double k, double log_tau,
                            int index_src);
/* src/titles.c:26-38 */
int findTitle(char **titles, const char *title, int ntitles);

/* include/fft_kernels.h:131-180: bilinear T(k, log tau) of one tabulated
 * function, DC -> 0; params: struct spline_params* */
void kernel_transfer_function(struct kernel *the_kernel);
void kernel_inv_transfer_function(struct kernel *the_kernel);

/* src/generate_grids.c:28-67: grid = c2r(T(k, log_tau_source) * grf) times
 * rescale_factor (the growth ratio) when it is not 1; the field file is
 * written by writeFieldFile_dg when fname is not NULL */
int generatePerturbationGrid(const struct cosmology *cosmo, const struct perturb_spline *spline,
                             struct distributed_grid *grf, struct distributed_grid *grid,
                             const char *transfer_func_title, const char *fname,
                             const double rescale_factor);

/* include/poisson.h:29-30 */
int solvePoisson(double *phi, double *f, int N, double boxlen); /* device N^3 arrays */
int solvePoisson_dg(struct distributed_grid *dg);

/* include/2lpt.h:27-29 */
int solve2LPT(struct distributed_grid *potential, struct distributed_grid *density,
              struct distributed_grid *workspace, double factor1, double factor2);

/* include/monge_ampere.h:27-29: cycles of the Monge-Ampere iteration
 * det(I + D^2 phi) = 1 + delta on top of the Zel'dovich potential; returns 1
 * for a density not in momentum space (src/monge_ampere.c:82-85) */
int solveMongeAmpere(struct distributed_grid *potential, struct distributed_grid *density,
                     struct distributed_grid *workspace, int cycles);
/* the root-mean-square residual of each cycle of this thread's last
 * solveMongeAmpere (the value src/monge_ampere.c:183-186 prints): copies up
 * to n values, returns the number of cycles */
int mitos_monge_ampere_residuals(double *rms_eps, int n);

/* include/calc_powerspec.h:26-36 -- box1/box2 device, bin arrays host.  The
 * reference returns void; a failure leaves NaN bins and mitos_last_error() set */
void calc_cross_powerspec(int N, double boxlen, const fftw_complex *box1,
                          const fftw_complex *box2, int bins, double *k_in_bins,
                          double *power_in_bins, int *obs_in_bins);
void calc_cross_powerspec_float(int N, double boxlen, const fftwf_complex *box1,
                                const fftwf_complex *box2, int bins, double *k_in_bins,
                                double *power_in_bins, int *obs_in_bins);
void calc_cross_powerspec_float_mpi(long int N, long int X0, long int NX, double boxlen,
                                    const fftwf_complex *box1, const fftwf_complex *box2,
                                    int bins, double *k_in_bins, double *power_in_bins,
                                    long int *obs_in_bins);

/* include/distributed_grid.h:62-72 */
struct left_right_slice {
    double *left_slice;  /* device, padded rows of N + 2 values */
    double *local_slice;
    double *right_slice;
    int left_NX;
    int left_X0;
    int local_NX;
    int local_X0;
    int right_NX;
    int right_X0;
};
/* include/grids_interp.h:36 -- one position (one small device launch; the
 * batched mitos_displace_velocity_tsc_dg is the form for particle loops);
 * NaN and mitos_last_error() set on a failure */
double gridTSC_dg(struct left_right_slice *lrs, double x, double y, double z, double boxlen,
                  int N);

/* include/grids_interp.h:41-48 -- device arrays */
int undoCICWindow(fftw_complex *farr, int N, double boxlen);
int undoTSCWindow(fftw_complex *farr, int N, double boxlen);
int undoCICWindowFloat(fftwf_complex *farr, int N, double boxlen);
int undoTSCWindowFloat(fftwf_complex *farr, int N, double boxlen);

/* ---- field files (HDF5: Header/BoxSize, Field/Field doubles) and shrinking */
/* libhdf5 is loaded at the first call (MITOS_HDF5_LIB, the default search
 * path, /opt/conda/lib); without it these return MITOS_ERR_UNSUPPORTED. */
/* include/output.h:31 -- host array N x N x N */
int writeFieldFile(const double *box, int N, double box_len, const char *fname);
/* include/input.h:149 -- allocates *box (malloc, host, N^3 unpadded) */
int readFieldFile(double **box, int *N, double *box_len, const char *fname);
/* include/output_mpi.h:37 -- the padded device box as an N x N x (N+2)
 * dataset; returns 1 for a grid in momentum space */
int writeFieldFile_dg(struct distributed_grid *dg, const char *fname);
/* include/input_mpi.h:27 -- a padded or unpadded dataset into the device box */
int readFieldFile_dg(struct distributed_grid *dg, const char *fname);
/* include/shrink_grids.h:26-29 -- M^3 block averages (M divides N, else 1);
 * out device M^3; shrinkGrid's in is a device N^3 array */
int shrinkGrid_dg(double *out, struct distributed_grid *in, int M, int N);
int shrinkGrid(double *out, const double *in, int M, int N);

/* The FFTW plans the reference builds inline (src/poisson.c:35-36,
 * analyse/src/mitos_read.c:411-413, analyse/src/mitos_dist_read.c:413):
 * unnormalised 3D transforms of unpadded N^3 device arrays. */
int mitos_fft_r2c(int N, const double *in, fftw_complex *out);
int mitos_fft_c2r(int N, const fftw_complex *in, double *out);
int mitos_fft_r2c_float(int N, const float *in, fftwf_complex *out);
int mitos_fft_c2r_float(int N, const fftwf_complex *in, float *out);

/* host-side table of the reference P(k) bin for every integer
 * n^2 = nx^2+ny^2+nz^2 in [0, 3(N/2)^2] (-1 for n^2 = 0); table has
 * 3(N/2)^2+1 entries.  No device work. */
int mitos_pk_bin_table(int N, double boxlen, int bins, int *table);

/* last error of a void reference function; mitos_strerror for any code */
int mitos_last_error(void);
const char *mitos_strerror(int code);
/* free the library's cached device scratch buffers */
int mitos_release_workspace(void);

/* ======================================================================== */
/* Part 2: batched particle steps (replace per-particle reference loops)     */
/* ======================================================================== */

/* src/particle.c:87-129 with FirstID (src/particle_types.c:102-104): M^3
 * lattice particles, X = x*((float)L/M) in float, id = x*M*M+y*M+z+first_id.
 * Device arrays; ids may be NULL. */
int mitos_lattice(int M, double boxlen, long long first_id, float *X, float *Y, float *Z,
                  long long *ids);

/* src/mitos.c:827-863 then :1072-1077: for dir = x, y, z in turn, subtract the
 * TSC-interpolated displacement grid (gridTSC_dg, src/grids_interp.c:267-309)
 * at the particle's current float position, then wrap into [0, L).  The
 * grids are in configuration space (padded box arrays). */
int mitos_displace_tsc_dg(const struct distributed_grid *psi_x,
                          const struct distributed_grid *psi_y,
                          const struct distributed_grid *psi_z, long long npart, float *X,
                          float *Y, float *Z);

/* src/mitos.c:824-908 then :1072-1077, the whole per-particle loop: the
 * three sequential displacement moves (psi grids, may all be NULL to skip),
 * then the velocities v_dir = (float)gridTSC_dg(vel_dir) at the moved, not yet
 * wrapped position (vel grids and VX/VY/VZ may all be NULL), then the wrap.
 * Configuration-space grids; device arrays. */
int mitos_displace_velocity_tsc_dg(const struct distributed_grid *psi_x,
                                   const struct distributed_grid *psi_y,
                                   const struct distributed_grid *psi_z,
                                   const struct distributed_grid *vel_x,
                                   const struct distributed_grid *vel_y,
                                   const struct distributed_grid *vel_z, long long npart,
                                   float *X, float *Y, float *Z, float *VX, float *VY, float *VZ);

/* src/mitos.c:1097-1107: the particle arrays as the IC writer lays them out
 * before H5Dwrite -- coords and vels npart x 3 doubles (x, y, z per
 * particle), masses npart doubles, ids npart (index + first_id).  Device
 * inputs (VX/VY/VZ may be NULL: zero velocities), host outputs (any may be
 * NULL). */
int mitos_particles_export(long long npart, const float *X, const float *Y, const float *Z,
                           const float *VX, const float *VY, const float *VZ, double mass,
                           long long first_id, double *coords, double *vels, double *masses,
                           long long *ids);

/* analyse/src/mitos_read.c:278-376: CIC mass assignment of npart particles of
 * equal mass into an unpadded N^3 device grid, then (box-avg)/avg.  The
 * deposit accumulates 2^-40-unit fixed point, so the result does not depend
 * on thread order. */
int mitos_cic_overdensity(int N, double boxlen, long long npart, const float *X,
                          const float *Y, const float *Z, double *box);

/* No counterpart in the reference: mitos_cic_overdensity for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  It
 * calls no allocation routine: the fixed-point grid is a fixed device array
 * of MITOS_NOALLOC_MAX_CELLS cells (N^3 <= 2^21, N <= 128; 16 MiB held on
 * every GPU the library runs on until the process ends), so larger grids
 * return MITOS_ERR_TOO_LARGE.  Work is issued on `stream` (a hipStream_t)
 * without a host wait; NULL issues it on the null stream and waits.  Calls
 * that overlap on one GPU share the array and race.  Same result, bit for
 * bit, as mitos_cic_overdensity. */
#define MITOS_NOALLOC_MAX_CELLS (1L << 21)
int mitos_cic_overdensity_noalloc(int N, double boxlen, long long npart, const float *X,
                                  const float *Y, const float *Z, double *box, void *stream);

/* ======================================================================== */
/* Part 3: fused pipeline (GRF -> 2LPT -> displacement -> CIC -> P(k))      */
/* ======================================================================== */

#define MITOS_FP64 0
#define MITOS_FP32 1

struct mitos_ic_config {
    int N;          /* [Box] GridSize */
    double boxlen;  /* [Box] BoxLen (Mpc/h) */
    uint64_t seed;  /* [Random] Seed (stream of rank 0, src/mitos.c:148) */
    double n_s;     /* [Cosmology] n_s */
    double sigma8;  /* [Spectrum] Sigma8 (top hat R = 8 Mpc/h) */
    double gamma;   /* [Spectrum] Gamma (BBKS shape) */
    double f1, f2;  /* 2LPT factors, (1, -3/7) as src/mitos.c:385 */
    int run2lpt;    /* [ParticleType_0] Run2LPT (0: Zel'dovich) */
    int bins;       /* [Read] PowerSpectrumBins */
    int precision;  /* MITOS_FP64 or MITOS_FP32 */
    int M;          /* [ParticleType_0] CubeRootNumber: M^3 lattice particles
                       (src/particle.c:87-129); 0 means M = N.  A context takes
                       at most the M^3 of the config it was created with;
                       ensembles and the noise forward/backward need M = N */
};

struct mitos_ic_result {
    /* host arrays of cfg.bins entries, caller-owned (may be NULL) */
    double *k_in_bins;
    double *power_in_bins;
    long long *obs_in_bins;
    double amplitude;   /* A of P(k) = A k^n_s T(k)^2 */
    float stage_ms[8];  /* noise, hessian, poisson, displacement, particles, cic, pk, total */
};

typedef struct mitos_ic_context mitos_ic_context;

/* sigma8 normalisation of P(k) = A k^n_s T_BBKS(k)^2 (host) */
double mitos_sigma8_amplitude(double n_s, double gamma, double sigma8, double R);

/* read the .ini keys of the path (see struct mitos_ic_config) */
int mitos_read_ini(const char *fname, struct mitos_ic_config *cfg);

int mitos_ic_create(const struct mitos_ic_config *cfg, mitos_ic_context **ctx);
/* run one realization; cfg must keep N and precision of the context */
int mitos_ic_run(mitos_ic_context *ctx, const struct mitos_ic_config *cfg,
                 struct mitos_ic_result *res);
/* the same, issued on `stream` (a hipStream_t; NULL: the context's stream)
 * without a host synchronisation */
int mitos_ic_enqueue(mitos_ic_context *ctx, const struct mitos_ic_config *cfg, void *stream);
/* P(k) of the last enqueued realization; this call, and the copies below,
 * first wait for the stream that realization was enqueued on */
int mitos_ic_fetch_pk(mitos_ic_context *ctx, struct mitos_ic_result *res);
/* copy the last run's particle positions (M^3 floats each, M^3 = N^3 unless
 * the config sets M) into caller device arrays */
int mitos_ic_positions(mitos_ic_context *ctx, float *X, float *Y, float *Z);
/* copy the last run's overdensity (unpadded N^3) or its three displacement
 * grids (3 x N^3, x then y then z, unpadded) into a caller device buffer of
 * the context's real type (float for MITOS_FP32, double for MITOS_FP64) */
int mitos_ic_density(mitos_ic_context *ctx, void *delta_device);
int mitos_ic_displacements(mitos_ic_context *ctx, void *psi_device);
/* particle ids (device, M^3 of the last run) written on demand, closed form
 * of the lattice */
int mitos_ic_ids(mitos_ic_context *ctx, long long *ids_device);
int mitos_ic_destroy(mitos_ic_context *ctx);

/* FFT schedule picked by the autotuner at mitos_ic_create for pass `kind`
 * (0: y forward, 1: x forward, 2: z c2r, 3: z r2c): returns the variant
 * (0..3) and fills ms[4] with the timed milliseconds of each variant (all 0
 * when MITOS_FFT_SCHEDULE=default disabled the timing); -1 on bad arguments */
int mitos_ic_schedule(mitos_ic_context *ctx, int kind, float *ms);
/* mean milliseconds of a plain 3D r2c and c2r of the context's size and
 * precision (real rows one half-complex row long, half-complex rows of
 * 128-byte multiples)
 * over reps runs, with the autotuned schedules; overwrites the context's
 * grids (run the pipeline again before reading its outputs) */
int mitos_ic_fft_timing(mitos_ic_context *ctx, int reps, float *r2c_ms, float *c2r_ms);
/* pin the schedule variants[4] (per kind as above) for contexts and ensembles
 * of size N and this precision created afterwards on the current device,
 * instead of timing them */
int mitos_set_fft_schedule(int N, int precision, const int *variants);

/* per-kernel profile of the following runs (HIP events on the run's stream):
 * mitos_ic_profile_entry returns the number of entries and fills entry i
 * (kernel name, milliseconds, algorithmic bytes) when 0 <= i < count */
int mitos_ic_profile(mitos_ic_context *ctx, int enable);
int mitos_ic_profile_entry(mitos_ic_context *ctx, int i, const char **name, float *ms,
                           double *bytes);

/* the reference's call order through the Part 1 functions (FP64 plain path);
 * X, Y, Z: optional device arrays of M^3 floats (M^3 = N^3 unless cfg->M is
 * set) that receive the particle positions, or NULL */
int mitos_ic_run_reference_order(const struct mitos_ic_config *cfg,
                                 struct mitos_ic_result *res, float *X, float *Y, float *Z);

/* ---- adjoint: gradient of a P(k) or density loss w.r.t. the white noise -- */
/* noise / grad_noise: device, dense N x N x (N/2+1) complex of the context's
 * real type (unit-variance white noise before the spectrum amplitude, in the
 * layout of generate_complex_grf's output with the row stride N/2+1).
 * _forward runs the pipeline from this noise (lattice TSC stencils applied in
 * k-space for both precisions, the form the adjoint differentiates) and
 * returns P(k) like mitos_ic_run.  _backward recomputes that forward run and
 * writes dL/dnoise (d/dRe + i d/dIm per mode) given dL/dP (host, cfg.bins
 * values; bins with no modes are ignored) and/or dL/ddelta (device, dense N^3
 * of the real type), either may be NULL.  The phi_,ij grids are recomputed,
 * not stored. */
int mitos_ic_noise_forward(mitos_ic_context *ctx, const struct mitos_ic_config *cfg,
                           const void *noise, struct mitos_ic_result *res);
int mitos_ic_noise_backward(mitos_ic_context *ctx, const struct mitos_ic_config *cfg,
                            const void *noise, const double *dL_dP, const void *dL_ddelta,
                            void *grad_noise);

/* ======================================================================== */
/* Part 4: ensembles -- n realizations of uneven size (any mix of N, boxlen, */
/* seed, n_s, bins, precision) captured once as one HIP graph and replayed.  */
/* Realizations of one (N, precision) reuse one set of grids in sequence;    */
/* groups of different sizes run concurrently.  Each realization's P(k) is   */
/* identical to mitos_ic_run with the same config.                           */
/* ======================================================================== */
typedef struct mitos_ensemble mitos_ensemble;

/* allocates the groups' grids, runs every realization once, captures the graph */
int mitos_ensemble_create(const struct mitos_ic_config *cfgs, int n, mitos_ensemble **ens);
/* one graph launch of all realizations; _run waits for it */
int mitos_ensemble_run(mitos_ensemble *ens);
int mitos_ensemble_enqueue(mitos_ensemble *ens, void *stream);
/* P(k) of realization i of the last launch */
int mitos_ensemble_fetch_pk(mitos_ensemble *ens, int i, struct mitos_ic_result *res);
int mitos_ensemble_destroy(mitos_ensemble *ens);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif

#endif /* MITOS_HIP_H */
