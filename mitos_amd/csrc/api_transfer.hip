// Tabulated transfer functions T(k, log tau): the host-side spline search of
// src/perturb_spline.c, the kernel_(inv_)transfer_function callbacks of
// include/fft_kernels.h:131-180 and generatePerturbationGrid of
// src/generate_grids.c:28-67.  Applied through fft_apply_kernel_dg, the
// callbacks run as one device pass (k_apply_transfer, kspace.h).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mitos_hip.h"
#include "kspace.h"

using namespace mitos;

// src/perturb_spline.c:30-64: the acceleration table holds, for k_acc_size
// evenly spaced v in [k_min, k_max), the largest index j with k_j < v
extern "C" int initPerturbSpline(struct perturb_spline* spline, int k_acc_size,
                                 const struct perturb_data* ptdat) {
    if (!spline || !ptdat || k_acc_size < 1 || ptdat->k_size < 2)
        return set_last_error(MITOS_ERR_ARG);
    spline->ptdat = ptdat;
    spline->k_acc_table = (double*)malloc(sizeof(double) * k_acc_size);
    spline->k_acc_table_size = k_acc_size;
    if (!spline->k_acc_table) return 1;
    const int n = ptdat->k_size;
    const double k_min = ptdat->k[0], k_max = ptdat->k[n - 1];
    for (int i = 0; i < k_acc_size; i++) {
        const double v = k_min + ((double)i / k_acc_size) * (k_max - k_min);
        // the table is increasing: the last index below v by bisection
        int lo = 0, hi = n - 1;
        if (!(ptdat->k[0] < v)) {
            spline->k_acc_table[i] = 0;
            continue;
        }
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (ptdat->k[mid] < v)
                lo = mid;
            else
                hi = mid - 1;
        }
        spline->k_acc_table[i] = lo;
    }
    return 0;
}

extern "C" int cleanPerturbSpline(struct perturb_spline* spline) {
    if (!spline) return set_last_error(MITOS_ERR_ARG);
    free(spline->k_acc_table);
    spline->k_acc_table = nullptr;
    return 0;
}

// src/perturb_spline.c:73-97
extern "C" int perturbSplineFindTau(const struct perturb_spline* spline, double log_tau,
                                    int* index, double* u) {
    const struct perturb_data* pt = spline->ptdat;
    const int n = pt->tau_size;
    if (log_tau < pt->log_tau[0]) {
        *index = 0;
        *u = 0.0;
        return 0;
    }
    if (log_tau >= pt->log_tau[n - 1]) {
        *index = n - 2;
        *u = 1.0;
        return 0;
    }
    // the first i >= 1 with log_tau_i >= log_tau, minus one
    int i = 1;
    while (i < n && pt->log_tau[i] < log_tau) i++;
    *index = i - 1;
    const double left = pt->log_tau[*index], right = pt->log_tau[*index + 1];
    *u = (log_tau - left) / (right - left);
    return 0;
}

// src/perturb_spline.c:100-138: the interval from the acceleration table's
// start index onwards
extern "C" int perturbSplineFindK(const struct perturb_spline* spline, double k, int* index,
                                  double* u) {
    const struct perturb_data* pt = spline->ptdat;
    const int n = pt->k_size;
    const double k_min = pt->k[0], k_max = pt->k[n - 1];
    if (k > k_max) {
        *index = n - 2;
        *u = 1.0;
        return 0;
    }
    int i = 0;  // below k_min (no defined result there): the first interval
    if (k >= k_min) {
        const int size = spline->k_acc_table_size;
        const double w = (k - k_min) / (k_max - k_min);
        int J = (int)floor(w * size);
        J = J < size ? J : size - 1;
        i = (int)spline->k_acc_table[J];
        while (i < n - 2 && !(k >= pt->k[i] && k <= pt->k[i + 1])) i++;
    }
    *index = i;
    const double left = pt->k[i], right = pt->k[i + 1];
    *u = (k - left) / (right - left);
    return 0;
}

// src/perturb_spline.c:141-160
extern "C" double perturbSplineInterp(const struct perturb_spline* spline, int k_index,
                                      int tau_index, double u_k, double u_tau, int index_src) {
#pragma clang fp contract(off)
    const struct perturb_data* pt = spline->ptdat;
    const int ks = pt->k_size, ts = pt->tau_size;
    const double* arr = pt->delta + (size_t)index_src * ks * ts;
    const double T11 = arr[(size_t)ks * tau_index + k_index];
    const double T21 = arr[(size_t)ks * tau_index + k_index + 1];
    const double T12 = arr[(size_t)ks * (tau_index + 1) + k_index];
    const double T22 = arr[(size_t)ks * (tau_index + 1) + k_index + 1];
    return (1 - u_tau) * ((1 - u_k) * T11 + u_k * T21) + u_tau * ((1 - u_k) * T12 + u_k * T22);
}

// src/perturb_spline.c:163-177
extern "C" double perturbSplineInterp0(const struct perturb_spline* spline, double k,
                                       double log_tau, int index_src) {
    int k_index = 0, tau_index = 0;
    double u_k, u_tau;
    perturbSplineFindTau(spline, log_tau, &tau_index, &u_tau);
    perturbSplineFindK(spline, k, &k_index, &u_k);
    return perturbSplineInterp(spline, k_index, tau_index, u_k, u_tau, index_src);
}

// src/titles.c:26-38: the last title of equal length and text
extern "C" int findTitle(char** titles, const char* title, int ntitles) {
    const size_t len = strlen(title);
    int found = -1;
    for (int i = 0; i < ntitles; i++) {
        if (strlen(titles[i]) != len) continue;
        if (strncmp(titles[i], title, len) == 0) found = i;
    }
    return found;
}

// include/fft_kernels.h:131-180 (host evaluation)
static double transfer_host(struct kernel* the_kernel) {
    const struct spline_params* sp = (const struct spline_params*)the_kernel->params;
    int k_index;
    double u_k;
    perturbSplineFindK(sp->spline, the_kernel->k, &k_index, &u_k);
    return perturbSplineInterp(sp->spline, k_index, sp->tau_index, u_k, sp->u_tau,
                               sp->index_src);
}

extern "C" void kernel_transfer_function(struct kernel* the_kernel) {
    double* kp = reinterpret_cast<double*>(&the_kernel->kern);
    kp[0] = the_kernel->k == 0 ? 0.0 : transfer_host(the_kernel);
    kp[1] = 0;
}

extern "C" void kernel_inv_transfer_function(struct kernel* the_kernel) {
    double* kp = reinterpret_cast<double*>(&the_kernel->kern);
    kp[0] = the_kernel->k == 0 ? 0.0 : 1.0 / transfer_host(the_kernel);
    kp[1] = 0;
}

// src/generate_grids.c:28-67
extern "C" int generatePerturbationGrid(const struct cosmology* cosmo,
                                        const struct perturb_spline* spline,
                                        struct distributed_grid* grf,
                                        struct distributed_grid* grid,
                                        const char* transfer_func_title, const char* fname,
                                        const double rescale_factor) {
    if (!cosmo || !spline || !spline->ptdat || !grf || !grid || !transfer_func_title)
        return set_last_error(MITOS_ERR_ARG);
    int tau_index;
    double u_tau;
    perturbSplineFindTau(spline, cosmo->log_tau_source, &tau_index, &u_tau);
    const int index_src =
        findTitle(spline->ptdat->titles, transfer_func_title, spline->ptdat->n_functions);
    if (index_src < 0) return set_last_error(1);  // src/generate_grids.c:44-47
    struct spline_params sp = {spline, index_src, tau_index, u_tau};
    MITOS_TRY(fft_apply_kernel_dg(grid, grf, kernel_transfer_function, &sp));
    if (rescale_factor != 1.0)
        MITOS_TRY(fft_apply_kernel_dg(grid, grid, kernel_constant, &rescale_factor));
    MITOS_TRY(fft_c2r_dg(grid));
    if (fname) return writeFieldFile_dg(grid, fname);
    return 0;
}
