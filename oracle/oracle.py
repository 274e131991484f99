"""ctypes front end of the FP64 CPU oracle (oracle/mitos_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package
(``mitos_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_mitos.so")
REF_XOSHIRO = os.path.join(HERE, "_ref", "xoshiro_ref")

_lib = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)


def build(force: bool = False) -> None:
    """Compile the oracle (and, when the reference tree exists, oracle/_ref)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "liboracle_mitos.so"], check=True,
                       stdout=subprocess.DEVNULL)
    if os.path.isdir("/root/reference/include"):
        subprocess.run(["make", "-C", HERE, "ref"], check=True, stdout=subprocess.DEVNULL)


class OrcGrid(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("boxlen", ctypes.c_double),
                ("momentum_space", ctypes.c_char), ("fbox", ctypes.c_void_p),
                ("box", ctypes.c_void_p)]


class OrcRng(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64 * 4)]


class OrcIcConfig(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("boxlen", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("n_s", ctypes.c_double), ("sigma8", ctypes.c_double),
                ("gamma", ctypes.c_double), ("f1", ctypes.c_double), ("f2", ctypes.c_double),
                ("run2lpt", ctypes.c_int), ("bins", ctypes.c_int), ("M", ctypes.c_int)]


class OrcIcOutputs(ctypes.Structure):
    _fields_ = [("psi", c_double_p), ("X", c_float_p), ("Y", c_float_p), ("Z", c_float_p),
                ("id", c_ll_p), ("density", c_double_p), ("k_in_bins", c_double_p),
                ("power_in_bins", c_double_p), ("obs_in_bins", c_int_p),
                ("delta_k", ctypes.c_void_p), ("stage_seconds", ctypes.c_double * 8)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_rng_init.restype = OrcRng
        L.orc_rng_init.argtypes = [ctypes.c_uint64]
        L.orc_rng_draws.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, c_u64_p]
        L.orc_rng_state_after.argtypes = [ctypes.c_uint64, ctypes.c_uint64, c_u64_p]
        L.orc_rng_draws_at.argtypes = [ctypes.c_uint64, c_u64_p, ctypes.c_uint64,
                                       ctypes.c_uint64, c_u64_p]
        L.orc_sample_norms.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                       c_double_p]
        L.orc_generate_complex_grf.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcRng)]
        L.orc_enforce_hermiticity.argtypes = [ctypes.POINTER(OrcGrid)]
        L.orc_fft_r2c.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_fft_c2r.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_apply_kernel.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, c_double_p]
        L.orc_bbks_transfer.restype = ctypes.c_double
        L.orc_bbks_transfer.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_sigma8_amplitude.restype = ctypes.c_double
        L.orc_sigma8_amplitude.argtypes = [ctypes.c_double] * 4
        L.orc_lattice.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_longlong, c_float_p,
                                  c_float_p, c_float_p, c_ll_p]
        L.orc_grid_tsc.restype = ctypes.c_double
        L.orc_grid_tsc.argtypes = [c_double_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double]
        L.orc_displace_tsc.argtypes = [ctypes.c_int, ctypes.c_double, c_double_p, c_double_p,
                                       c_double_p, ctypes.c_longlong, c_float_p, c_float_p,
                                       c_float_p]
        L.orc_cic.restype = ctypes.c_double
        L.orc_cic.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_longlong, c_float_p,
                              c_float_p, c_float_p, ctypes.c_double, c_double_p]
        L.orc_overdensity.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      c_double_p]
        L.orc_powerspec.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, c_double_p, c_double_p,
                                    c_int_p]
        L.orc_solve_2lpt.argtypes = [ctypes.POINTER(OrcGrid)] * 3 + [ctypes.c_double] * 2
        L.orc_apply_kernel_dg.argtypes = [ctypes.POINTER(OrcGrid), ctypes.POINTER(OrcGrid),
                                          ctypes.c_int, c_double_p]
        L.orc_fft_r2c_dg.argtypes = [ctypes.POINTER(OrcGrid)]
        L.orc_fft_c2r_dg.argtypes = [ctypes.POINTER(OrcGrid)]
        L.orc_ic_run.argtypes = [ctypes.POINTER(OrcIcConfig), ctypes.POINTER(OrcIcOutputs)]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_fft_threads.restype = ctypes.c_int
        L.orc_set_fft_threads.argtypes = [ctypes.c_int]
        L.orc_pk_bin.restype = ctypes.c_int
        L.orc_pk_bin.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int]
        L.orc_pk_lut_mismatches.restype = ctypes.c_longlong
        L.orc_pk_lut_mismatches.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                            c_int_p]
        L.orc_spline_acc_table.argtypes = [c_double_p, ctypes.c_int, ctypes.c_int, c_double_p]
        L.orc_spline_find_tau.argtypes = [c_double_p, ctypes.c_int, ctypes.c_double, c_int_p,
                                          c_double_p]
        L.orc_spline_find_k.argtypes = [c_double_p, ctypes.c_int, c_double_p, ctypes.c_int,
                                        ctypes.c_double, c_int_p, c_double_p]
        L.orc_spline_interp.restype = ctypes.c_double
        L.orc_spline_interp.argtypes = [c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int]
        L.orc_apply_transfer.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_double, c_double_p, ctypes.c_int, c_double_p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_displace_velocity_tsc.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_longlong, c_float_p,
                                                c_float_p, c_float_p, c_float_p, c_float_p,
                                                c_float_p]
        L.orc_solve_monge_ampere.argtypes = [ctypes.POINTER(OrcGrid)] * 3 + [ctypes.c_int,
                                                                            c_double_p]
        L.orc_ranlxd_draws.argtypes = [ctypes.c_ulong, ctypes.c_long, c_double_p]
        L.orc_generate_ngeniclike_grf.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_long]
        L.orc_shrink_grid.argtypes = [c_double_p, c_double_p, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


# kernel ids (mitos_oracle.h)
K_CONSTANT, K_INV_POISSON, K_DX, K_DY, K_DZ = 0, 1, 2, 3, 4
K_POWER_NO_TRANSFER, K_UNDO_WINDOW, K_SQRT_POWER_LAW, K_BBKS = 5, 6, 7, 8
K_LOWPASS, K_HIPASS, K_TOPHAT, K_ELLIPTIC_TOPHAT = 9, 10, 11, 12
K_TRANSLATE, K_GAUSSIAN, K_GAUSSIAN_INV = 13, 14, 15


class Transfer:
    """A table of transfer functions T_f(k, log tau) in the layout of
    struct perturb_data (include/perturb_data.h:25-52): delta[f][tau][k]."""

    def __init__(self, k, log_tau, delta, acc_size=1000):
        self.k = np.ascontiguousarray(k, dtype=np.float64)
        self.log_tau = np.ascontiguousarray(log_tau, dtype=np.float64)
        self.delta = np.ascontiguousarray(delta, dtype=np.float64)
        assert self.delta.shape[1:] == (self.log_tau.size, self.k.size)
        self.acc_size = acc_size
        self.acc = np.empty(acc_size)
        lib().orc_spline_acc_table(_p(self.k, c_double_p), self.k.size, acc_size,
                                   _p(self.acc, c_double_p))

    def find_tau(self, lt):
        i = ctypes.c_int()
        u = ctypes.c_double()
        lib().orc_spline_find_tau(_p(self.log_tau, c_double_p), self.log_tau.size, lt,
                                  ctypes.byref(i), ctypes.byref(u))
        return i.value, u.value

    def find_k(self, kk):
        i = ctypes.c_int()
        u = ctypes.c_double()
        lib().orc_spline_find_k(_p(self.k, c_double_p), self.k.size, _p(self.acc, c_double_p),
                                self.acc_size, kk, ctypes.byref(i), ctypes.byref(u))
        return i.value, u.value

    def interp(self, ki, ti, uk, ut, f):
        return lib().orc_spline_interp(_p(self.delta, c_double_p), self.k.size,
                                       self.log_tau.size, ki, ti, uk, ut, f)

    def apply(self, fbox, boxlen, f, tau_index, u_tau, invert=False):
        N = fbox.shape[0]
        r = np.ascontiguousarray(fbox, dtype=np.complex128)
        w = np.empty_like(r)
        lib().orc_apply_transfer(w.ctypes.data, r.ctypes.data, N, boxlen,
                                 _p(self.k, c_double_p), self.k.size,
                                 _p(self.delta, c_double_p), self.log_tau.size, f, tau_index,
                                 u_tau, self.acc_size, int(invert))
        return w


def ranlxd_draws(seed: int, n: int) -> np.ndarray:
    out = np.empty(n)
    lib().orc_ranlxd_draws(seed, n, _p(out, c_double_p))
    return out


def ngeniclike_grf(N: int, boxlen: float, seed: int) -> np.ndarray:
    out = np.zeros((N, N, N // 2 + 1), np.complex128)
    assert lib().orc_generate_ngeniclike_grf(out.ctypes.data, N, boxlen, seed) == 0
    return out


def shrink_grid(box: np.ndarray, M: int) -> np.ndarray:
    N = box.shape[0]
    b = np.ascontiguousarray(box, dtype=np.float64)
    out = np.empty((M, M, M))
    assert lib().orc_shrink_grid(_p(out, c_double_p), _p(b, c_double_p), M, N) == 0
    return out


def set_threads(n: int) -> int:
    return lib().orc_set_threads(n)


def draws(seed: int, skip: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().orc_rng_draws(seed, skip, n, _p(out, c_u64_p))
    return out


def draws_at(seed: int, offsets, k: int) -> np.ndarray:
    """k draws at each of the given draw offsets (one pass over the stream);
    returns shape (len(offsets), k) in the order given"""
    offs = np.asarray(offsets, dtype=np.uint64)
    order = np.argsort(offs, kind="stable")
    srt = np.ascontiguousarray(offs[order])
    out = np.empty((len(offs), k), dtype=np.uint64)
    tmp = np.empty((len(offs), k), dtype=np.uint64)
    lib().orc_rng_draws_at(seed, _p(srt, c_u64_p), len(offs), k, _p(tmp, c_u64_p))
    out[order] = tmp
    return out


def state_after(seed: int, n: int) -> np.ndarray:
    out = np.empty(4, dtype=np.uint64)
    lib().orc_rng_state_after(seed, n, _p(out, c_u64_p))
    return out


def normals(seed: int, skip_draws: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_sample_norms(seed, skip_draws, n, _p(out, c_double_p))
    return out


def ref_draws(seed: int, skip: int, n: int) -> np.ndarray:
    """Draws from the reference's own random_xorshift.h (oracle/_ref build)."""
    out = subprocess.run([REF_XOSHIRO, str(seed), str(skip), str(n)], check=True,
                         capture_output=True, text=True).stdout.split()
    return np.array([int(v) for v in out], dtype=np.uint64)


def grf(N: int, boxlen: float, seed: int, hermitian: bool = True) -> np.ndarray:
    """src/grf.c:26-69 (+ :78-175): complex white noise times sqrt(L^3/2)."""
    L = lib()
    fbox = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
    box = np.zeros((N, N, N + 2), dtype=np.float64)
    g = OrcGrid(N, boxlen, b"\x00", fbox.ctypes.data, box.ctypes.data)
    st = L.orc_rng_init(seed)
    L.orc_generate_complex_grf(ctypes.byref(g), ctypes.byref(st))
    if hermitian:
        L.orc_enforce_hermiticity(ctypes.byref(g))
    return fbox


def fft_r2c(box: np.ndarray) -> np.ndarray:
    """Unnormalised 3D r2c of an (N,N,N) or padded (N,N,N+2) real array."""
    N = box.shape[0]
    box = np.ascontiguousarray(box, dtype=np.float64)
    out = np.empty((N, N, N // 2 + 1), dtype=np.complex128)
    lib().orc_fft_r2c(N, box.ctypes.data, box.shape[2], out.ctypes.data)
    return out


def fft_c2r(fbox: np.ndarray, padded: bool = False) -> np.ndarray:
    """Unnormalised 3D c2r (FFTW semantics) to (N,N,N) or padded (N,N,N+2)."""
    N = fbox.shape[0]
    fbox = np.ascontiguousarray(fbox, dtype=np.complex128)
    rs = N + 2 if padded else N
    out = np.zeros((N, N, rs), dtype=np.float64)
    lib().orc_fft_c2r(N, fbox.ctypes.data, out.ctypes.data, rs)
    return out


def apply_kernel(fbox: np.ndarray, boxlen: float, kid: int, params=()) -> np.ndarray:
    N = fbox.shape[0]
    r = np.ascontiguousarray(fbox, dtype=np.complex128)
    w = np.empty_like(r)
    prm = np.array(list(params) + [0.0] * 4, dtype=np.float64)
    lib().orc_apply_kernel(w.ctypes.data, r.ctypes.data, N, boxlen, kid, _p(prm, c_double_p))
    return w


def bbks(k: float, gamma: float) -> float:
    return lib().orc_bbks_transfer(k, gamma)


def sigma8_amplitude(n_s: float, gamma: float, sigma8: float, R: float = 8.0) -> float:
    return lib().orc_sigma8_amplitude(n_s, gamma, sigma8, R)


def lattice(M: int, boxlen: float, first_id: int = 1):
    n = M ** 3
    X = np.empty(n, np.float32)
    Y = np.empty(n, np.float32)
    Z = np.empty(n, np.float32)
    ids = np.empty(n, np.int64)
    lib().orc_lattice(M, boxlen, first_id, _p(X, c_float_p), _p(Y, c_float_p),
                      _p(Z, c_float_p), _p(ids, c_ll_p))
    return X, Y, Z, ids


def grid_tsc(box_padded: np.ndarray, boxlen: float, x: float, y: float, z: float) -> float:
    N = box_padded.shape[0]
    b = np.ascontiguousarray(box_padded, dtype=np.float64)
    return lib().orc_grid_tsc(_p(b, c_double_p), N, boxlen, x, y, z)


def displace_tsc(psi: np.ndarray, boxlen: float, X, Y, Z):
    """psi: (3, N, N, N+2) padded displacement grids. Positions updated in place."""
    N = psi.shape[1]
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    lib().orc_displace_tsc(N, boxlen, _p(psi[0], c_double_p), _p(psi[1], c_double_p),
                           _p(psi[2], c_double_p), X.size, _p(X, c_float_p),
                           _p(Y, c_float_p), _p(Z, c_float_p))


def displace_velocity_tsc(psi, vel, boxlen: float, X, Y, Z):
    """psi, vel: (3, N, N, N+2) padded grids or None.  Positions updated in
    place; returns the velocities (three float arrays) or None."""
    N = (psi if psi is not None else vel).shape[1]
    n = X.size
    keep = []

    def ptrs(g):
        if g is None:
            return None
        g = np.ascontiguousarray(g, dtype=np.float64)
        keep.append(g)
        arr = (ctypes.c_void_p * 3)(*[g[i].ctypes.data for i in range(3)])
        keep.append(arr)
        return ctypes.cast(arr, ctypes.c_void_p)
    V = [np.zeros(n, np.float32) for _ in range(3)]
    lib().orc_displace_velocity_tsc(N, boxlen, ptrs(psi), ptrs(vel), n, _p(X, c_float_p),
                                    _p(Y, c_float_p), _p(Z, c_float_p), _p(V[0], c_float_p),
                                    _p(V[1], c_float_p), _p(V[2], c_float_p))
    return V if vel is not None else None


def cic(N: int, boxlen: float, X, Y, Z, mass: float = 1.0):
    box = np.zeros((N, N, N), dtype=np.float64)
    tot = lib().orc_cic(N, boxlen, X.size, _p(X, c_float_p), _p(Y, c_float_p),
                        _p(Z, c_float_p), mass, _p(box, c_double_p))
    return box, tot


def overdensity(box: np.ndarray, boxlen: float, total_mass: float) -> None:
    lib().orc_overdensity(box.shape[0], boxlen, total_mass, _p(box, c_double_p))


def powerspec(fbox1: np.ndarray, fbox2: np.ndarray, boxlen: float, bins: int):
    N = fbox1.shape[0]
    a = np.ascontiguousarray(fbox1, dtype=np.complex128)
    b = np.ascontiguousarray(fbox2, dtype=np.complex128)
    k = np.empty(bins)
    P = np.empty(bins)
    obs = np.empty(bins, dtype=np.int32)
    with np.errstate(all="ignore"):
        lib().orc_powerspec(N, boxlen, a.ctypes.data, b.ctypes.data, bins, _p(k, c_double_p),
                            _p(P, c_double_p), _p(obs, c_int_p))
    return k, P, obs


def pk_bin(N: int, boxlen: float, bins: int, x: int, y: int, z: int) -> int:
    return lib().orc_pk_bin(N, boxlen, bins, x, y, z)


def pk_lut_mismatches(N: int, boxlen: float, bins: int, lut: np.ndarray) -> int:
    lut = np.ascontiguousarray(lut, dtype=np.int32)
    return lib().orc_pk_lut_mismatches(N, boxlen, bins, _p(lut, c_int_p))


class Grid:
    """Host-memory stand-in of struct distributed_grid for oracle calls."""

    def __init__(self, N: int, boxlen: float):
        self.fbox = np.zeros((N, N, N // 2 + 1), dtype=np.complex128)
        self.box = np.zeros((N, N, N + 2), dtype=np.float64)
        self.s = OrcGrid(N, boxlen, b"\x00", self.fbox.ctypes.data, self.box.ctypes.data)

    @property
    def momentum_space(self) -> int:
        return ord(self.s.momentum_space)

    def ref(self):
        return ctypes.byref(self.s)


def solve_2lpt(pot: Grid, dens: Grid, work: Grid, f1: float, f2: float) -> int:
    return lib().orc_solve_2lpt(pot.ref(), dens.ref(), work.ref(), f1, f2)


def solve_monge_ampere(pot: Grid, dens: Grid, work: Grid, cycles: int):
    eps = np.zeros(max(cycles, 1))
    rc = lib().orc_solve_monge_ampere(pot.ref(), dens.ref(), work.ref(), cycles,
                                      _p(eps, c_double_p))
    return rc, eps[:cycles]


def ic_run(N: int, boxlen: float, seed: int, n_s: float = 0.965, sigma8: float = 0.81,
           gamma: float = 0.21, bins: int = 32, run2lpt: bool = True, f1: float = 1.0,
           f2: float = -3.0 / 7.0, want_psi: bool = False, want_density: bool = False,
           want_delta_k: bool = False, threads: int | None = None,
           fft_threads: int | None = None, M: int = 0):
    """Reference-order FP64 pipeline; returns a dict of numpy outputs.
    threads: OpenMP threads of the loops; fft_threads: of the FFTs alone
    (1 = serial, as the reference's FFTW plans; None = as the loops)."""
    L = lib()
    if threads is not None:
        L.orc_set_threads(threads)
    L.orc_set_fft_threads(fft_threads or 0)
    n = (M or N) ** 3  # lattice particles
    cfg = OrcIcConfig(N, boxlen, seed, n_s, sigma8, gamma, f1, f2, int(run2lpt), bins, M)
    res = {"X": np.empty(n, np.float32), "Y": np.empty(n, np.float32),
           "Z": np.empty(n, np.float32), "id": np.empty(n, np.int64),
           "k": np.empty(bins), "P": np.empty(bins), "obs": np.empty(bins, np.int32)}
    out = OrcIcOutputs()
    out.X = _p(res["X"], c_float_p)
    out.Y = _p(res["Y"], c_float_p)
    out.Z = _p(res["Z"], c_float_p)
    out.id = _p(res["id"], c_ll_p)
    out.k_in_bins = _p(res["k"], c_double_p)
    out.power_in_bins = _p(res["P"], c_double_p)
    out.obs_in_bins = _p(res["obs"], c_int_p)
    if want_psi:
        res["psi"] = np.zeros((3, N, N, N + 2))
        out.psi = _p(res["psi"], c_double_p)
    if want_density:
        res["density"] = np.zeros((N, N, N))
        out.density = _p(res["density"], c_double_p)
    if want_delta_k:
        res["delta_k"] = np.zeros((N, N, N // 2 + 1), np.complex128)
        out.delta_k = res["delta_k"].ctypes.data
    with np.errstate(all="ignore"):
        err = L.orc_ic_run(ctypes.byref(cfg), ctypes.byref(out))
    if err:
        raise RuntimeError(f"oracle pipeline failed with code {err}")
    res["stage_seconds"] = list(out.stage_seconds)[:7]
    return res
