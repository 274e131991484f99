/*
 * fft_kernels.h -- replacement of the reference's include/fft_kernels.h for
 * translation units linked against libmitos_hip.so.
 *
 * The reference defines every k-space kernel callback as a static inline
 * function in this header (include/fft_kernels.h:28-219), so each unit that
 * includes include/mitos.h passes the address of its own copy to
 * fft_apply_kernel_dg.  libmitos_hip.so evaluates such unknown callbacks on
 * the host, mode by mode.  Putting this file in place of the reference's
 * include/fft_kernels.h makes the same names refer to the library's exported
 * callbacks, whose addresses select the device kernels: the calls of
 * src/mitos.c, src/2lpt.c, src/poisson.c, src/generate_grids.c and the
 * analyse tools then run on the GPU without other source changes.
 *
 * It is meant to be included where the reference includes its own
 * fft_kernels.h (from include/mitos.h, after fft.h, with the reference's
 * primordial.h and perturb_spline.h beside it), so struct kernel, struct
 * cosmology and struct perturb_spline come from the reference headers.
 * Called directly, each callback evaluates the kernel on the host.
 */
#ifndef FFT_KERNELS_H
#define FFT_KERNELS_H

#include <math.h>

/* in a reference unit the types come from the reference's headers; after
 * mitos_hip.h (which declares the same types) nothing more is needed */
#ifndef MITOS_HIP_H
#include "primordial.h"
#include "perturb_spline.h"
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*kernel_func)(struct kernel *the_kernel);

/* include/fft_kernels.h:28-93 */
void kernel_constant(struct kernel *the_kernel);        /* params: double          */
void kernel_lowpass(struct kernel *the_kernel);         /* params: double k_max    */
void kernel_hipass(struct kernel *the_kernel);          /* params: double k_min    */
void kernel_tophat(struct kernel *the_kernel);          /* params: double[2]       */
void kernel_elliptic_tophat(struct kernel *the_kernel); /* params: double[4]       */
void kernel_translate(struct kernel *the_kernel);       /* params: double[3]       */
void kernel_gaussian(struct kernel *the_kernel);        /* params: double R        */
void kernel_gaussian_inv(struct kernel *the_kernel);    /* params: double R        */

/* include/fft_kernels.h:96-121 */
void kernel_inv_poisson(struct kernel *the_kernel);
void kernel_dx(struct kernel *the_kernel);
void kernel_dy(struct kernel *the_kernel);
void kernel_dz(struct kernel *the_kernel);
void kernel_power_no_transfer(struct kernel *the_kernel); /* params: struct cosmology* */

/* include/fft_kernels.h:123-180 */
#ifndef MITOS_HIP_H
struct spline_params {
    const struct perturb_spline *spline;
    int index_src;
    int tau_index;
    double u_tau;
};
#endif
void kernel_transfer_function(struct kernel *the_kernel);     /* params: spline_params* */
void kernel_inv_transfer_function(struct kernel *the_kernel); /* params: spline_params* */

/* include/fft_kernels.h:187 */
static inline double sinc(double x) { return x == 0 ? 1. : sin(x) / x; }

/* include/fft_kernels.h:189-219 */
#ifndef MITOS_HIP_H
struct Hermite_kern_params {
    int order;
    int N;
    double boxlen;
};
#endif
void kernel_undo_Hermite_window(struct kernel *the_kernel); /* params: Hermite_kern_params* */

#ifdef __cplusplus
}
#endif

#endif /* FFT_KERNELS_H */
