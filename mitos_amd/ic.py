"""Python mirror of the reference's IC call surface over libmitos_hip.so.

Names follow the reference (``alloc_local_grid``, ``generate_complex_grf``,
``fft_r2c_dg``, ``solve2LPT``, ``calc_cross_powerspec`` ...) with the same
argument meaning and return codes; grids live in GPU memory.  Device arrays
passed to the serial-array functions are torch CUDA tensors.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import MITOS_FP32, MITOS_FP64, MitosError, check, kernel_ptr


def lib():
    return _lib.load()


def _sync():
    import torch
    torch.cuda.synchronize()


def _ptr(t) -> int:
    return t.data_ptr()


class Grid:
    """A ``struct distributed_grid`` on one rank (NX = N, X0 = 0) in device memory."""

    def __init__(self, N: int, boxlen: float):
        self.s = _lib.DistributedGrid()
        check(lib().alloc_local_grid(ctypes.byref(self.s), N, boxlen, 0x44000000),
              "alloc_local_grid")

    @property
    def N(self) -> int:
        return self.s.N

    @property
    def boxlen(self) -> float:
        return self.s.boxlen

    @property
    def momentum_space(self) -> int:
        return ord(self.s.momentum_space)

    @momentum_space.setter
    def momentum_space(self, v: int) -> None:
        self.s.momentum_space = bytes([v])

    def ref(self):
        return ctypes.byref(self.s)

    def fbox(self) -> np.ndarray:
        N = self.N
        out = np.empty((N, N, N // 2 + 1), np.complex128)
        check(lib().mitos_grid_to_host(self.ref(), out.ctypes.data, None), "mitos_grid_to_host")
        return out

    def box(self) -> np.ndarray:
        N = self.N
        out = np.empty((N, N, N + 2), np.float64)
        check(lib().mitos_grid_to_host(self.ref(), None, out.ctypes.data), "mitos_grid_to_host")
        return out

    def set_fbox(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, np.complex128)
        check(lib().mitos_grid_from_host(self.ref(), a.ctypes.data, None), "mitos_grid_from_host")

    def set_box(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, np.float64)
        check(lib().mitos_grid_from_host(self.ref(), None, a.ctypes.data), "mitos_grid_from_host")

    def free(self) -> None:
        if self.s.fbox or self.s.box:
            lib().free_local_grid(self.ref())

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rand_uint64_init(seed: int):
    return lib().rand_uint64_init(seed)


def generate_complex_grf(g: Grid, state) -> int:
    return lib().generate_complex_grf(g.ref(), ctypes.byref(state))


def enforce_hermiticity(g: Grid) -> int:
    return lib().enforce_hermiticity(g.ref())


def fft_r2c_dg(g: Grid) -> int:
    return lib().fft_r2c_dg(g.ref())


def fft_c2r_dg(g: Grid) -> int:
    return lib().fft_c2r_dg(g.ref())


def kernel_params(name: str, **kw):
    """ctypes parameter block for a kernel callback (kept alive by the caller)."""
    if name == "kernel_constant":
        return ctypes.c_double(kw["constant"])
    if name == "kernel_power_no_transfer":
        c = _lib.Cosmology()
        c.A_s, c.n_s, c.k_pivot = kw["A_s"], kw["n_s"], kw["k_pivot"]
        return c
    if name == "kernel_undo_Hermite_window":
        return _lib.HermiteKernParams(kw["order"], kw["N"], kw["boxlen"])
    if name in ("kernel_sqrt_power_law", "kernel_bbks_transfer"):
        return _lib.SpectrumParams(kw.get("A", 0.0), kw.get("n_s", 0.0), kw.get("Gamma", 0.0))
    if name in ("kernel_lowpass", "kernel_hipass", "kernel_gaussian", "kernel_gaussian_inv"):
        return ctypes.c_double(kw["value"])
    if name in ("kernel_tophat", "kernel_elliptic_tophat", "kernel_translate"):
        v = list(kw["values"])
        return (ctypes.c_double * len(v))(*v)
    return None


class PerturbTable:
    """Transfer functions T_f(k, log tau) as ``struct perturb_data`` with its
    ``struct perturb_spline`` (initPerturbSpline); the numpy arrays are kept
    alive with the structs.  delta has shape (n_functions, tau_size, k_size)."""

    def __init__(self, k, log_tau, delta, titles, acc_size: int = 1000):
        self.k = np.ascontiguousarray(k, np.float64)
        self.log_tau = np.ascontiguousarray(log_tau, np.float64)
        self.delta = np.ascontiguousarray(delta, np.float64)
        nf, nt, nk = self.delta.shape
        assert nt == self.log_tau.size and nk == self.k.size and len(titles) == nf
        self._titles = [t.encode() for t in titles]
        self.titles = (ctypes.c_char_p * nf)(*self._titles)
        dp = lambda a: a.ctypes.data_as(_lib.c_double_p)
        self.data = _lib.PerturbData(nk, nt, nf, dp(self.delta), dp(self.k), dp(self.log_tau),
                                     None, None, None, None, self.titles, None)
        self.spline = _lib.PerturbSpline()
        check(lib().initPerturbSpline(ctypes.byref(self.spline), acc_size,
                                      ctypes.byref(self.data)), "initPerturbSpline")

    def params(self, index_src: int, log_tau: float) -> "_lib.SplineParams":
        """struct spline_params for kernel_(inv_)transfer_function at log_tau"""
        ti = ctypes.c_int()
        ut = ctypes.c_double()
        lib().perturbSplineFindTau(ctypes.byref(self.spline), log_tau, ctypes.byref(ti),
                                   ctypes.byref(ut))
        return _lib.SplineParams(ctypes.pointer(self.spline), index_src, ti.value, ut.value)

    def close(self) -> None:
        if self.spline.k_acc_table:
            lib().cleanPerturbSpline(ctypes.byref(self.spline))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generatePerturbationGrid(cosmo, table: PerturbTable, grf: Grid, grid: Grid, title: str,
                             fname: str | None = None, rescale: float = 1.0) -> int:
    return lib().generatePerturbationGrid(ctypes.byref(cosmo), ctypes.byref(table.spline),
                                          grf.ref(), grid.ref(), title.encode(),
                                          fname.encode() if fname else None, rescale)


def fft_apply_kernel_dg(w: Grid, r: Grid, name: str, params=None) -> int:
    p = ctypes.cast(ctypes.byref(params), ctypes.c_void_p) if params is not None else None
    return lib().fft_apply_kernel_dg(w.ref(), r.ref(), kernel_ptr(name), p)


def solvePoisson_dg(g: Grid) -> int:
    return lib().solvePoisson_dg(g.ref())


def solve2LPT(pot: Grid, dens: Grid, work: Grid, f1: float, f2: float) -> int:
    return lib().solve2LPT(pot.ref(), dens.ref(), work.ref(), f1, f2)


def calc_cross_powerspec(N: int, boxlen: float, box1, box2, bins: int, fp32: bool = False):
    """box1/box2: torch CUDA tensors in the (N, N, N/2+1) complex layout."""
    _sync()
    k = np.empty(bins)
    P = np.empty(bins)
    obs = np.empty(bins, np.int32)
    fn = lib().calc_cross_powerspec_float if fp32 else lib().calc_cross_powerspec
    fn(N, boxlen, _ptr(box1), _ptr(box2), bins, k.ctypes.data_as(_lib.c_double_p),
       P.ctypes.data_as(_lib.c_double_p), obs.ctypes.data_as(_lib.c_int_p))
    check(lib().mitos_last_error(), "calc_cross_powerspec")
    return k, P, obs


def sigma8_amplitude(n_s: float, gamma: float, sigma8: float, R: float = 8.0) -> float:
    return lib().mitos_sigma8_amplitude(n_s, gamma, sigma8, R)


def pk_bin_table(N: int, boxlen: float, bins: int) -> np.ndarray:
    t = np.empty(3 * (N // 2) ** 2 + 1, np.int32)
    check(lib().mitos_pk_bin_table(N, boxlen, bins, t.ctypes.data), "mitos_pk_bin_table")
    return t


@dataclass
class ICConfig:
    """struct mitos_ic_config (the .ini keys of the path)."""
    N: int = 64
    boxlen: float = 256.0
    seed: int = 42
    n_s: float = 0.965
    sigma8: float = 0.81
    gamma: float = 0.21
    f1: float = 1.0
    f2: float = -3.0 / 7.0
    run2lpt: int = 1
    bins: int = 32
    precision: int = MITOS_FP64
    M: int = 0  # particle lattice M^3 (CubeRootNumber); 0 means M = N

    @property
    def npart(self) -> int:
        return (self.M or self.N) ** 3

    def c(self) -> _lib.IcConfig:
        return _lib.IcConfig(self.N, self.boxlen, self.seed, self.n_s, self.sigma8, self.gamma,
                             self.f1, self.f2, self.run2lpt, self.bins, self.precision, self.M)

    @staticmethod
    def from_ini(path: str) -> "ICConfig":
        c = _lib.IcConfig()
        check(lib().mitos_read_ini(path.encode(), ctypes.byref(c)), "mitos_read_ini")
        return ICConfig(c.N, c.boxlen, c.seed, c.n_s, c.sigma8, c.gamma, c.f1, c.f2, c.run2lpt,
                        c.bins, c.precision, c.M)


class IC:
    """A fused-pipeline context (mitos_ic_create / run / destroy)."""

    def __init__(self, cfg: ICConfig):
        self.cfg = cfg
        self.h = ctypes.c_void_p()
        check(lib().mitos_ic_create(ctypes.byref(cfg.c()), ctypes.byref(self.h)),
              "mitos_ic_create")

    def run(self, cfg: ICConfig | None = None) -> dict:
        cfg = cfg or self.cfg
        bins = cfg.bins
        k = np.empty(bins)
        P = np.empty(bins)
        obs = np.empty(bins, np.int64)
        res = _lib.IcResult(k.ctypes.data_as(_lib.c_double_p), P.ctypes.data_as(_lib.c_double_p),
                            obs.ctypes.data_as(_lib.c_ll_p), 0.0)
        check(lib().mitos_ic_run(self.h, ctypes.byref(cfg.c()), ctypes.byref(res)),
              "mitos_ic_run")
        self._npart = cfg.npart
        return {"k": k, "P": P, "obs": obs, "amplitude": res.amplitude,
                "stage_ms": list(res.stage_ms)}

    def enqueue(self, cfg: ICConfig | None = None, stream: int | None = None) -> None:
        cfg = cfg or self.cfg
        check(lib().mitos_ic_enqueue(self.h, ctypes.byref(cfg.c()), stream), "mitos_ic_enqueue")
        self._npart = cfg.npart

    def fetch_pk(self) -> dict:
        """P(k) of the last enqueued realization (waits for its stream)."""
        bins = self.cfg.bins
        k = np.empty(bins)
        P = np.empty(bins)
        obs = np.empty(bins, np.int64)
        res = _lib.IcResult(k.ctypes.data_as(_lib.c_double_p), P.ctypes.data_as(_lib.c_double_p),
                            obs.ctypes.data_as(_lib.c_ll_p), 0.0)
        check(lib().mitos_ic_fetch_pk(self.h, ctypes.byref(res)), "mitos_ic_fetch_pk")
        return {"k": k, "P": P, "obs": obs, "amplitude": res.amplitude}

    def noise_forward(self, cfg: ICConfig, noise) -> dict:
        """P(k) of the pipeline run from `noise` (torch CUDA complex tensor of
        shape (N, N, N/2+1), complex64 for FP32 contexts, complex128 for FP64)."""
        self._check_noise(cfg, noise)
        bins = cfg.bins
        k = np.empty(bins)
        P = np.empty(bins)
        obs = np.empty(bins, np.int64)
        res = _lib.IcResult(k.ctypes.data_as(_lib.c_double_p), P.ctypes.data_as(_lib.c_double_p),
                            obs.ctypes.data_as(_lib.c_ll_p), 0.0)
        _sync()
        check(lib().mitos_ic_noise_forward(self.h, ctypes.byref(cfg.c()), _ptr(noise),
                                           ctypes.byref(res)), "mitos_ic_noise_forward")
        return {"k": k, "P": P, "obs": obs, "amplitude": res.amplitude}

    def noise_backward(self, cfg: ICConfig, noise, dL_dP=None, dL_ddelta=None):
        """dL/dnoise (same shape and dtype as `noise`) for a loss with gradient
        dL_dP (bins values) and/or dL_ddelta (CUDA tensor (N, N, N))."""
        import torch
        self._check_noise(cfg, noise)
        grad = torch.empty_like(noise)
        gp = None
        if dL_dP is not None:
            gp = np.ascontiguousarray(np.nan_to_num(np.asarray(dL_dP, np.float64)))
            assert gp.size == cfg.bins
        gd = None
        if dL_ddelta is not None:
            gd = dL_ddelta.contiguous().to(self._real())
            assert gd.numel() == cfg.N ** 3 and gd.is_cuda
        _sync()
        check(lib().mitos_ic_noise_backward(
            self.h, ctypes.byref(cfg.c()), _ptr(noise),
            gp.ctypes.data_as(_lib.c_double_p) if gp is not None else None,
            _ptr(gd) if gd is not None else None, _ptr(grad)), "mitos_ic_noise_backward")
        return grad

    def _check_noise(self, cfg, noise):
        import torch
        N = cfg.N
        want = torch.complex64 if self.cfg.precision == MITOS_FP32 else torch.complex128
        if (tuple(noise.shape) != (N, N, N // 2 + 1) or noise.dtype != want or not noise.is_cuda
                or not noise.is_contiguous()):
            raise ValueError(f"noise must be a contiguous CUDA {want} tensor of shape "
                             f"({N}, {N}, {N // 2 + 1})")

    def fft_timing(self, reps: int = 5) -> tuple:
        """(r2c ms, c2r ms) of a plain 3D transform of this size (overwrites the
        context's grids)."""
        a = ctypes.c_float()
        b = ctypes.c_float()
        check(lib().mitos_ic_fft_timing(self.h, reps, ctypes.byref(a), ctypes.byref(b)),
              "mitos_ic_fft_timing")
        return a.value, b.value

    def schedules(self) -> dict:
        """Autotuned FFT schedule per pass kind: (variant, [ms per variant])."""
        out = {}
        for kind, name in enumerate(("y_fwd", "x_fwd", "z_c2r", "z_r2c")):
            ms = (ctypes.c_float * 4)()
            v = lib().mitos_ic_schedule(self.h, kind, ms)
            out[name] = (v, [float(x) for x in ms])
        return out

    def profile(self, enable: bool = True) -> None:
        check(lib().mitos_ic_profile(self.h, int(enable)), "mitos_ic_profile")

    def profile_entries(self) -> list:
        """[(kernel, ms, algorithmic bytes)] of the last profiled run."""
        name = ctypes.c_char_p()
        ms = ctypes.c_float()
        by = ctypes.c_double()
        n = lib().mitos_ic_profile_entry(self.h, -1, None, None, None)
        out = []
        for i in range(n):
            lib().mitos_ic_profile_entry(self.h, i, ctypes.byref(name), ctypes.byref(ms),
                                         ctypes.byref(by))
            out.append((name.value.decode(), ms.value, by.value))
        return out

    def positions(self):
        import torch
        n = getattr(self, "_npart", self.cfg.npart)
        X = torch.empty(n, dtype=torch.float32, device="cuda")
        Y = torch.empty_like(X)
        Z = torch.empty_like(X)
        _sync()
        check(lib().mitos_ic_positions(self.h, _ptr(X), _ptr(Y), _ptr(Z)), "mitos_ic_positions")
        return X, Y, Z

    def _real(self):
        import torch
        return torch.float32 if self.cfg.precision == MITOS_FP32 else torch.float64

    def displacements(self):
        import torch
        N = self.cfg.N
        out = torch.empty((3, N, N, N), dtype=self._real(), device="cuda")
        _sync()
        check(lib().mitos_ic_displacements(self.h, _ptr(out)), "mitos_ic_displacements")
        return out

    def density(self):
        import torch
        N = self.cfg.N
        out = torch.empty((N, N, N), dtype=self._real(), device="cuda")
        _sync()
        check(lib().mitos_ic_density(self.h, _ptr(out)), "mitos_ic_density")
        return out

    def ids(self):
        import torch
        out = torch.empty(getattr(self, "_npart", self.cfg.npart), dtype=torch.int64,
                          device="cuda")
        _sync()
        check(lib().mitos_ic_ids(self.h, _ptr(out)), "mitos_ic_ids")
        return out

    def close(self) -> None:
        if self.h:
            lib().mitos_ic_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_fft_schedule(N: int, precision: int, variants) -> None:
    """Pin the FFT schedule variants (y fwd, x fwd, z c2r, z r2c) of later
    contexts of this size and precision (mitos_set_fft_schedule)."""
    arr = (ctypes.c_int * 4)(*variants)
    check(lib().mitos_set_fft_schedule(N, precision, arr), "mitos_set_fft_schedule")


def run_reference_order(cfg: ICConfig, positions: bool = False) -> dict:
    """The reference's call order through the drop-in functions (FP64 plain path)."""
    bins = cfg.bins
    k = np.empty(bins)
    P = np.empty(bins)
    obs = np.empty(bins, np.int64)
    res = _lib.IcResult(k.ctypes.data_as(_lib.c_double_p), P.ctypes.data_as(_lib.c_double_p),
                        obs.ctypes.data_as(_lib.c_ll_p), 0.0)
    X = Y = Z = None
    ptrs = (None, None, None)
    if positions:
        import torch
        n = cfg.npart
        X = torch.empty(n, dtype=torch.float32, device="cuda")
        Y = torch.empty_like(X)
        Z = torch.empty_like(X)
        _sync()
        ptrs = (_ptr(X), _ptr(Y), _ptr(Z))
    c = cfg.c()
    c.precision = MITOS_FP64
    check(lib().mitos_ic_run_reference_order(ctypes.byref(c), ctypes.byref(res), *ptrs),
          "mitos_ic_run_reference_order")
    out = {"k": k, "P": P, "obs": obs, "amplitude": res.amplitude,
           "stage_ms": list(res.stage_ms)}
    if positions:
        out.update(X=X, Y=Y, Z=Z)
    return out


class Ensemble:
    """n realizations of uneven size as one captured HIP graph
    (mitos_ensemble_create / run / fetch_pk / destroy)."""

    def __init__(self, cfgs: list):
        self.cfgs = list(cfgs)
        arr = (_lib.IcConfig * len(self.cfgs))(*[c.c() for c in self.cfgs])
        self.h = ctypes.c_void_p()
        check(lib().mitos_ensemble_create(arr, len(self.cfgs), ctypes.byref(self.h)),
              "mitos_ensemble_create")

    def run(self) -> None:
        check(lib().mitos_ensemble_run(self.h), "mitos_ensemble_run")

    def enqueue(self, stream: int | None = None) -> None:
        check(lib().mitos_ensemble_enqueue(self.h, stream), "mitos_ensemble_enqueue")

    def pk(self, i: int) -> dict:
        bins = self.cfgs[i].bins
        k = np.empty(bins)
        P = np.empty(bins)
        obs = np.empty(bins, np.int64)
        res = _lib.IcResult(k.ctypes.data_as(_lib.c_double_p), P.ctypes.data_as(_lib.c_double_p),
                            obs.ctypes.data_as(_lib.c_ll_p), 0.0)
        check(lib().mitos_ensemble_fetch_pk(self.h, i, ctypes.byref(res)),
              "mitos_ensemble_fetch_pk")
        return {"k": k, "P": P, "obs": obs, "amplitude": res.amplitude}

    def close(self) -> None:
        if self.h:
            lib().mitos_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ensemble_configs(scale: int = 1) -> list:
    """The 96-realization ensemble: 32 each at 64^3/250, 128^3/500 and
    256^3/1000 Mpc/h, seeds 1..96, n_s uniform over [0.94, 0.99] in seed
    order, FP32, 32 bins.  `scale` divides every N (smaller test ensembles)."""
    out = []
    for i in range(96):
        g = i // 32
        N = (64, 128, 256)[g] // scale
        L = (250.0, 500.0, 1000.0)[g]
        out.append(ICConfig(N=N, boxlen=L, seed=i + 1, n_s=0.94 + 0.05 * i / 95, bins=32,
                            run2lpt=1, precision=MITOS_FP32))
    return out


def white_noise(N: int, seed: int, precision: int = MITOS_FP64):
    """Unit-variance complex white noise of the seed's stream (the draws of
    generate_complex_grf, before any amplitude), as a CUDA tensor of shape
    (N, N, N/2+1) for the noise_* entry points."""
    import torch
    g = Grid(N, 2.0 ** (1.0 / 3.0))  # sqrt(L^3/2) = 1
    st = rand_uint64_init(seed)
    check(generate_complex_grf(g, st), "generate_complex_grf")
    out = torch.from_numpy(g.fbox()).cuda()
    g.free()
    return out.to(torch.complex64) if precision == MITOS_FP32 else out
