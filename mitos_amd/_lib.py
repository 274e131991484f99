"""ctypes bindings of libmitos_hip.so (include/mitos_hip.h).

The shared library is built in-tree by ``python -m mitos_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmitos_hip.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)
c_long_p = ctypes.POINTER(ctypes.c_long)
vp = ctypes.c_void_p

MITOS_FP64 = 0
MITOS_FP32 = 1

ERRORS = {
    0: "MITOS_OK", 1: "MITOS_ERR_DIMS", 2: "MITOS_ERR_NOT_MOMENTUM", 3: "MITOS_ERR_DEVICE",
    4: "MITOS_ERR_UNSUPPORTED", 5: "MITOS_ERR_ARG", 6: "MITOS_ERR_ALLOC",
    7: "MITOS_ERR_REAL_SPACE", 8: "MITOS_ERR_TOO_LARGE",
}


class MitosError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        super().__init__(f"{where} failed: {ERRORS.get(code, code)}")


class DistributedGrid(ctypes.Structure):
    """include/distributed_grid.h:29-60 (fbox/box are device pointers)."""
    _fields_ = [("N", ctypes.c_int), ("boxlen", ctypes.c_double), ("comm", ctypes.c_int),
                ("momentum_space", ctypes.c_char), ("NX", ctypes.c_long), ("X0", ctypes.c_long),
                ("local_size", ctypes.c_long), ("fbox", vp), ("box", vp)]


class Kernel(ctypes.Structure):
    """include/fft.h:31-39"""
    _fields_ = [("kx", ctypes.c_double), ("ky", ctypes.c_double), ("kz", ctypes.c_double),
                ("k", ctypes.c_double), ("kern", ctypes.c_double * 2), ("params", vp)]


class Cosmology(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("h", "n_s", "A_s", "k_pivot", "z_ini", "log_tau_ini", "rho_crit", "z_source",
                 "log_tau_source")]


class HermiteKernParams(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("N", ctypes.c_int), ("boxlen", ctypes.c_double)]


class SpectrumParams(ctypes.Structure):
    _fields_ = [("A", ctypes.c_double), ("n_s", ctypes.c_double), ("Gamma", ctypes.c_double)]


class RngState(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64 * 4)]


class IcConfig(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("boxlen", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("n_s", ctypes.c_double), ("sigma8", ctypes.c_double),
                ("gamma", ctypes.c_double), ("f1", ctypes.c_double), ("f2", ctypes.c_double),
                ("run2lpt", ctypes.c_int), ("bins", ctypes.c_int), ("precision", ctypes.c_int),
                ("M", ctypes.c_int)]


class IcResult(ctypes.Structure):
    _fields_ = [("k_in_bins", c_double_p), ("power_in_bins", c_double_p),
                ("obs_in_bins", c_ll_p), ("amplitude", ctypes.c_double),
                ("stage_ms", ctypes.c_float * 8)]


class PerturbData(ctypes.Structure):
    """include/perturb_data.h:25-52 (host arrays)"""
    _fields_ = [("k_size", ctypes.c_int), ("tau_size", ctypes.c_int),
                ("n_functions", ctypes.c_int), ("delta", c_double_p), ("k", c_double_p),
                ("log_tau", c_double_p), ("redshift", c_double_p), ("D_growth", c_double_p),
                ("f_growth", c_double_p), ("H_Hubble", c_double_p),
                ("titles", ctypes.POINTER(ctypes.c_char_p)), ("Omega", c_double_p)]


class PerturbSpline(ctypes.Structure):
    """include/perturb_spline.h:25-34"""
    _fields_ = [("ptdat", ctypes.POINTER(PerturbData)), ("k_acc_table", c_double_p),
                ("k_acc_table_size", ctypes.c_int)]


class SplineParams(ctypes.Structure):
    """include/fft_kernels.h:123-128"""
    _fields_ = [("spline", ctypes.POINTER(PerturbSpline)), ("index_src", ctypes.c_int),
                ("tau_index", ctypes.c_int), ("u_tau", ctypes.c_double)]


class LeftRightSlice(ctypes.Structure):
    """include/distributed_grid.h:62-72 (device slices)"""
    _fields_ = [("left_slice", vp), ("local_slice", vp), ("right_slice", vp),
                ("left_NX", ctypes.c_int), ("left_X0", ctypes.c_int),
                ("local_NX", ctypes.c_int), ("local_X0", ctypes.c_int),
                ("right_NX", ctypes.c_int), ("right_X0", ctypes.c_int)]


KERNEL_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(Kernel))
DG_P = ctypes.POINTER(DistributedGrid)

# (name, restype, argtypes) of every exported function in include/mitos_hip.h
SIGNATURES = [
    ("alloc_local_grid", ctypes.c_int, [ctypes.POINTER(DistributedGrid), ctypes.c_int,
                                        ctypes.c_double, ctypes.c_int]),
    ("free_local_grid", ctypes.c_int, [ctypes.POINTER(DistributedGrid)]),
    ("free_local_real_grid", ctypes.c_int, [ctypes.POINTER(DistributedGrid)]),
    ("free_local_complex_grid", ctypes.c_int, [ctypes.POINTER(DistributedGrid)]),
    ("mitos_grid_to_host", ctypes.c_int, [ctypes.POINTER(DistributedGrid), vp, vp]),
    ("mitos_grid_from_host", ctypes.c_int, [ctypes.POINTER(DistributedGrid), vp, vp]),
    ("rand_uint64", ctypes.c_uint64, [ctypes.POINTER(RngState)]),
    ("rand_uint64_init", RngState, [ctypes.c_uint64]),
    ("sampleUniform", ctypes.c_double, [ctypes.POINTER(RngState)]),
    ("sampleNorm", ctypes.c_double, [ctypes.POINTER(RngState)]),
    ("generate_complex_grf", ctypes.c_int, [ctypes.POINTER(DistributedGrid),
                                            ctypes.POINTER(RngState)]),
    ("enforce_hermiticity", ctypes.c_int, [ctypes.POINTER(DistributedGrid)]),
    ("generate_ngeniclike_grf", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_long, ctypes.c_double, ctypes.c_long]),
    ("fft_wavevector", None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_double, c_double_p, c_double_p, c_double_p, c_double_p]),
    ("fft_normalize_r2c", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
    ("fft_normalize_r2c_float", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
    ("fft_normalize_r2c_float_mpi", ctypes.c_int, [vp, ctypes.c_long, ctypes.c_long,
                                                   ctypes.c_double]),
    ("fft_normalize_c2r", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
    ("fft_apply_kernel", ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_double, vp, vp]),
    ("fft_apply_kernel_float", ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_double, vp, vp]),
    ("fft_apply_kernel_float_mpi", ctypes.c_int, [vp, vp, ctypes.c_long, ctypes.c_long,
                                                  ctypes.c_long, ctypes.c_double, vp, vp]),
    ("fft_r2c_dg", ctypes.c_int, [ctypes.POINTER(DistributedGrid)]),
    ("fft_c2r_dg", ctypes.c_int, [ctypes.POINTER(DistributedGrid)]),
    ("fft_apply_kernel_dg", ctypes.c_int, [ctypes.POINTER(DistributedGrid),
                                           ctypes.POINTER(DistributedGrid), vp, vp]),
    ("kernel_constant", None, [ctypes.POINTER(Kernel)]),
    ("kernel_lowpass", None, [ctypes.POINTER(Kernel)]),
    ("kernel_hipass", None, [ctypes.POINTER(Kernel)]),
    ("kernel_tophat", None, [ctypes.POINTER(Kernel)]),
    ("kernel_elliptic_tophat", None, [ctypes.POINTER(Kernel)]),
    ("kernel_translate", None, [ctypes.POINTER(Kernel)]),
    ("kernel_gaussian", None, [ctypes.POINTER(Kernel)]),
    ("kernel_gaussian_inv", None, [ctypes.POINTER(Kernel)]),
    ("kernel_inv_poisson", None, [ctypes.POINTER(Kernel)]),
    ("kernel_dx", None, [ctypes.POINTER(Kernel)]),
    ("kernel_dy", None, [ctypes.POINTER(Kernel)]),
    ("kernel_dz", None, [ctypes.POINTER(Kernel)]),
    ("kernel_power_no_transfer", None, [ctypes.POINTER(Kernel)]),
    ("kernel_undo_Hermite_window", None, [ctypes.POINTER(Kernel)]),
    ("kernel_sqrt_power_law", None, [ctypes.POINTER(Kernel)]),
    ("kernel_bbks_transfer", None, [ctypes.POINTER(Kernel)]),
    ("initPerturbSpline", ctypes.c_int, [ctypes.POINTER(PerturbSpline), ctypes.c_int,
                                         ctypes.POINTER(PerturbData)]),
    ("cleanPerturbSpline", ctypes.c_int, [ctypes.POINTER(PerturbSpline)]),
    ("perturbSplineFindTau", ctypes.c_int, [ctypes.POINTER(PerturbSpline), ctypes.c_double,
                                            c_int_p, c_double_p]),
    ("perturbSplineFindK", ctypes.c_int, [ctypes.POINTER(PerturbSpline), ctypes.c_double,
                                          c_int_p, c_double_p]),
    ("perturbSplineInterp", ctypes.c_double, [ctypes.POINTER(PerturbSpline), ctypes.c_int,
                                              ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_int]),
    ("perturbSplineInterp0", ctypes.c_double, [ctypes.POINTER(PerturbSpline), ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int]),
    ("findTitle", ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p,
                                 ctypes.c_int]),
    ("kernel_transfer_function", None, [ctypes.POINTER(Kernel)]),
    ("kernel_inv_transfer_function", None, [ctypes.POINTER(Kernel)]),
    ("generatePerturbationGrid", ctypes.c_int, [ctypes.POINTER(Cosmology),
                                                ctypes.POINTER(PerturbSpline), DG_P, DG_P,
                                                ctypes.c_char_p, ctypes.c_char_p,
                                                ctypes.c_double]),
    ("solvePoisson", ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_double]),
    ("solvePoisson_dg", ctypes.c_int, [ctypes.POINTER(DistributedGrid)]),
    ("solve2LPT", ctypes.c_int, [ctypes.POINTER(DistributedGrid)] * 3 + [ctypes.c_double] * 2),
    ("solveMongeAmpere", ctypes.c_int, [DG_P] * 3 + [ctypes.c_int]),
    ("mitos_monge_ampere_residuals", ctypes.c_int, [c_double_p, ctypes.c_int]),
    ("calc_cross_powerspec", None, [ctypes.c_int, ctypes.c_double, vp, vp, ctypes.c_int,
                                    c_double_p, c_double_p, c_int_p]),
    ("calc_cross_powerspec_float", None, [ctypes.c_int, ctypes.c_double, vp, vp, ctypes.c_int,
                                          c_double_p, c_double_p, c_int_p]),
    ("calc_cross_powerspec_float_mpi", None, [ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                              ctypes.c_double, vp, vp, ctypes.c_int,
                                              c_double_p, c_double_p, c_long_p]),
    ("gridTSC_dg", ctypes.c_double, [ctypes.POINTER(LeftRightSlice), ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int]),
    ("undoCICWindow", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
    ("undoTSCWindow", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
    ("undoCICWindowFloat", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
    ("undoTSCWindowFloat", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
    ("writeFieldFile", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_char_p]),
    ("readFieldFile", ctypes.c_int, [ctypes.POINTER(c_double_p), c_int_p, c_double_p,
                                     ctypes.c_char_p]),
    ("writeFieldFile_dg", ctypes.c_int, [DG_P, ctypes.c_char_p]),
    ("readFieldFile_dg", ctypes.c_int, [DG_P, ctypes.c_char_p]),
    ("shrinkGrid_dg", ctypes.c_int, [vp, DG_P, ctypes.c_int, ctypes.c_int]),
    ("shrinkGrid", ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int]),
    ("mitos_fft_r2c", ctypes.c_int, [ctypes.c_int, vp, vp]),
    ("mitos_fft_c2r", ctypes.c_int, [ctypes.c_int, vp, vp]),
    ("mitos_fft_r2c_float", ctypes.c_int, [ctypes.c_int, vp, vp]),
    ("mitos_fft_c2r_float", ctypes.c_int, [ctypes.c_int, vp, vp]),
    ("mitos_pk_bin_table", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, vp]),
    ("mitos_last_error", ctypes.c_int, []),
    ("mitos_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("mitos_release_workspace", ctypes.c_int, []),
    ("mitos_lattice", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_longlong, vp, vp,
                                     vp, vp]),
    ("mitos_displace_tsc_dg", ctypes.c_int, [ctypes.POINTER(DistributedGrid)] * 3 +
     [ctypes.c_longlong, vp, vp, vp]),
    ("mitos_displace_velocity_tsc_dg", ctypes.c_int, [DG_P] * 6 + [ctypes.c_longlong] +
     [vp] * 6),
    ("mitos_particles_export", ctypes.c_int, [ctypes.c_longlong] + [vp] * 6 +
     [ctypes.c_double, ctypes.c_longlong] + [vp] * 4),
    ("mitos_cic_overdensity_noalloc", ctypes.c_int, [ctypes.c_int, ctypes.c_double,
                                                     ctypes.c_longlong, vp, vp, vp, vp, vp]),
    ("mitos_cic_overdensity", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_longlong,
                                             vp, vp, vp, vp]),
    ("mitos_sigma8_amplitude", ctypes.c_double, [ctypes.c_double] * 4),
    ("mitos_read_ini", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(IcConfig)]),
    ("mitos_ic_create", ctypes.c_int, [ctypes.POINTER(IcConfig), ctypes.POINTER(vp)]),
    ("mitos_ic_run", ctypes.c_int, [vp, ctypes.POINTER(IcConfig), ctypes.POINTER(IcResult)]),
    ("mitos_ic_enqueue", ctypes.c_int, [vp, ctypes.POINTER(IcConfig), vp]),
    ("mitos_ic_fetch_pk", ctypes.c_int, [vp, ctypes.POINTER(IcResult)]),
    ("mitos_ic_positions", ctypes.c_int, [vp, vp, vp, vp]),
    ("mitos_ic_density", ctypes.c_int, [vp, vp]),
    ("mitos_ic_displacements", ctypes.c_int, [vp, vp]),
    ("mitos_ic_ids", ctypes.c_int, [vp, vp]),
    ("mitos_ic_destroy", ctypes.c_int, [vp]),
    ("mitos_ic_schedule", ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    ("mitos_ic_fft_timing", ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float)]),
    ("mitos_set_fft_schedule", ctypes.c_int, [ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
    ("mitos_ic_profile", ctypes.c_int, [vp, ctypes.c_int]),
    ("mitos_ic_profile_entry", ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                              ctypes.POINTER(ctypes.c_float),
                                              ctypes.POINTER(ctypes.c_double)]),
    ("mitos_ic_run_reference_order", ctypes.c_int, [ctypes.POINTER(IcConfig),
                                                    ctypes.POINTER(IcResult), vp, vp, vp]),
    ("mitos_ic_noise_forward", ctypes.c_int, [vp, ctypes.POINTER(IcConfig), vp,
                                              ctypes.POINTER(IcResult)]),
    ("mitos_ic_noise_backward", ctypes.c_int, [vp, ctypes.POINTER(IcConfig), vp, c_double_p, vp,
                                               vp]),
    ("mitos_ensemble_create", ctypes.c_int, [ctypes.POINTER(IcConfig), ctypes.c_int,
                                             ctypes.POINTER(vp)]),
    ("mitos_ensemble_run", ctypes.c_int, [vp]),
    ("mitos_ensemble_enqueue", ctypes.c_int, [vp, vp]),
    ("mitos_ensemble_fetch_pk", ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(IcResult)]),
    ("mitos_ensemble_destroy", ctypes.c_int, [vp]),
]

_lib = None


def load():
    """Load libmitos_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mitos_amd.build` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int, where: str) -> None:
    if code != 0:
        raise MitosError(code, where)


def kernel_ptr(name: str) -> int:
    """Address of an exported kernel callback (selects the device kernel)."""
    return ctypes.cast(getattr(load(), name), vp).value
