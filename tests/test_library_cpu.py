"""The C-ABI library on the host (CPU only, no GPU): it loads, exports every
function include/mitos_hip.h declares, and its host-side functions agree with
the oracle and closed forms.  No device compute is called here."""
import ctypes
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mitos_hip.h")


@pytest.fixture(scope="module")
def L():
    from mitos_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libmitos_hip.so is not built (python -m mitos_amd.build)")
    return _lib.load()


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    names = []
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", src, flags=re.M):
        head = src[m.start():m.end()]
        if head.lstrip().startswith(("typedef", "struct", "#", "return", "if")):
            continue
        names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for n in ("alloc_local_grid", "free_local_grid", "generate_complex_grf", "enforce_hermiticity",
              "fft_r2c_dg", "fft_c2r_dg", "fft_apply_kernel_dg", "fft_wavevector",
              "solvePoisson_dg", "solve2LPT", "calc_cross_powerspec",
              "calc_cross_powerspec_float", "undoCICWindow", "undoTSCWindow",
              "rand_uint64_init", "kernel_inv_poisson", "kernel_dx"):
        assert n in names, n


def test_library_exports_every_declared_function(L):
    from mitos_amd import _lib
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_exports_are_only_the_declared_ones(L):
    from mitos_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    ours = {n for n in exported if not n.startswith(("_", "__"))}
    assert ours == set(declared_functions()), ours ^ set(declared_functions())


def test_rng_matches_reference_stream(L, oracle):
    from mitos_amd import _lib
    with np.load(os.path.join(ROOT, "tests", "golden", "xoshiro.npz")) as f:
        ref = f["first42"][:64]
    st = L.rand_uint64_init(42)
    got = [L.rand_uint64(ctypes.byref(st)) for _ in range(64)]
    assert np.array_equal(np.array(got, dtype=np.uint64), ref)
    st = L.rand_uint64_init(7)
    assert [int(v) for v in st.s] == [int(v) for v in oracle.state_after(7, 0)]


def test_fft_wavevector(L):
    N, L_ = 16, 100.0
    dk = 2 * math.pi / L_
    out = [ctypes.c_double() for _ in range(4)]
    for x, y, z in ((0, 0, 0), (3, 9, 8), (8, 8, 8), (15, 1, 2)):
        L.fft_wavevector(x, y, z, N, dk, *[ctypes.byref(o) for o in out])
        f = lambda i: (i - N if i > N // 2 else i) * dk  # N/2 -> +N/2 (src/fft.c:31-33)
        kx, ky, kz = f(x), f(y), f(z)
        assert [o.value for o in out] == [kx, ky, kz, math.sqrt(kx * kx + ky * ky + kz * kz)]


def test_kernel_callbacks_on_host(L):
    from mitos_amd import _lib
    kx, ky, kz = 0.3, -0.2, 0.5
    k = math.sqrt(kx * kx + ky * ky + kz * kz)
    K = _lib.Kernel(kx, ky, kz, k)

    def ev(name, params=None):
        K.kern[0] = K.kern[1] = float("nan")
        K.params = ctypes.cast(ctypes.byref(params), ctypes.c_void_p) if params is not None \
            else None
        getattr(L, name)(ctypes.byref(K))
        return complex(K.kern[0], K.kern[1])

    c = ctypes.c_double(2.5)
    assert ev("kernel_constant", c) == 2.5
    assert ev("kernel_inv_poisson") == -1.0 / k / k
    assert ev("kernel_dx") == 1j * kx
    assert ev("kernel_dy") == 1j * ky
    assert ev("kernel_dz") == 1j * kz
    sp = _lib.SpectrumParams(3.0, 0.965, 0.21)
    assert abs(ev("kernel_sqrt_power_law", sp) - math.sqrt(3.0 * k ** 0.965)) < 1e-15
    q = k / 0.21
    T = math.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 +
                                               (6.71 * q) ** 4) ** -0.25
    assert abs(ev("kernel_bbks_transfer", sp) - T) < 1e-15
    h = _lib.HermiteKernParams(3, 64, 100.0)
    s = lambda x: math.sin(x) / x
    W = (s(kx * 100 / 128) * s(ky * 100 / 128) * s(kz * 100 / 128)) ** 3
    assert abs(ev("kernel_undo_Hermite_window", h) - 1 / W) < 1e-14
    cosmo = _lib.Cosmology()
    cosmo.A_s, cosmo.n_s, cosmo.k_pivot = 2.1e-9, 0.965, 0.05
    exp = math.sqrt(2.1e-9 * (k / 0.05) ** (0.965 - 1) * k * 2 * math.pi ** 2)
    assert abs(ev("kernel_power_no_transfer", cosmo) - exp) < 1e-15 * exp
    K.k = 0.0
    assert ev("kernel_inv_poisson") == 1.0  # DC passes through (include/fft_kernels.h:98)


def test_pk_bin_table_matches_oracle(L, oracle):
    # the reference bins every mode with its own float formula
    # (src/calc_powerspec.c:58-63); the library's table over n^2 must agree for
    # every mode: the configurations, the ensemble's members, radix-3 sizes,
    # odd box lengths and bin counts
    for N, Lb, bins in ((64, 256.0, 32), (512, 1000.0, 128), (1024, 2000.0, 128),
                        (64, 250.0, 32), (128, 500.0, 32), (256, 1000.0, 32),
                        (24, 96.0, 12), (48, 150.0, 16), (96, 300.0, 24), (384, 750.0, 64),
                        (32, 100.0 / 3.0, 16), (128, 64.0, 17), (256, 500.0, 200),
                        (1536, 2000.0, 128)):
        t = np.empty(3 * (N // 2) ** 2 + 1, np.int32)
        assert L.mitos_pk_bin_table(N, Lb, bins, t.ctypes.data) == 0
        assert oracle.pk_lut_mismatches(N, Lb, bins, t) == 0
        assert t[0] == -1 and t[1:].min() == 0 and t.max() == bins - 1
        assert np.all(np.diff(t[1:]) >= 0)  # nondecreasing in n^2


def test_sigma8_amplitude_matches_oracle(L, oracle):
    a = L.mitos_sigma8_amplitude(0.965, 0.21, 0.81, 8.0)
    assert abs(a / oracle.sigma8_amplitude(0.965, 0.21, 0.81, 8.0) - 1) < 1e-13


def test_read_ini(L, tmp_path):
    from mitos_amd import _lib, ic
    p = tmp_path / "ic.ini"
    p.write_text("""; parameters of the IC path
[Random]
Seed = 42   ; the seed
[Box]
GridSize = 64
BoxLen = 256.0
[Cosmology]
n_s = 0.965
[Spectrum]
Sigma8 = 0.81
Gamma = "0.21"
[ParticleType_0]
CubeRootNumber = 64
Run2LPT = 1
[Read]
PowerSpectrumBins = 32
[GPU]
Precision = FP32
""")
    c = _lib.IcConfig()
    assert L.mitos_read_ini(str(p).encode(), ctypes.byref(c)) == 0
    assert (c.N, c.boxlen, c.seed, c.n_s, c.sigma8, c.gamma, c.run2lpt, c.bins, c.precision) == \
        (64, 256.0, 42, 0.965, 0.81, 0.21, 1, 32, _lib.MITOS_FP32)
    assert (c.f1, c.f2) == (1.0, -3.0 / 7.0)
    assert c.M == 0  # CubeRootNumber = GridSize: M = N
    p.write_text("[Box]\nGridSize = 64\n[ParticleType_0]\nCubeRootNumber = 32\n")
    assert L.mitos_read_ini(str(p).encode(), ctypes.byref(c)) == 0
    assert (c.N, c.M) == (64, 32)
    assert ic.ICConfig.from_ini(str(p)).npart == 32 ** 3
    assert L.mitos_read_ini(b"/nonexistent.ini", ctypes.byref(c)) == 5


def test_error_strings(L):
    for code in range(8):
        s = L.mitos_strerror(code)
        assert s and len(s) > 3
    assert L.mitos_strerror(0) != L.mitos_strerror(3)


def test_python_front_end_config_roundtrip():
    import mitos_amd.ic as ic
    cfg = ic.ICConfig(N=128, boxlen=500.0, seed=9, bins=16, precision=ic.MITOS_FP32)
    c = cfg.c()
    assert (c.N, c.boxlen, c.seed, c.bins, c.precision) == (128, 500.0, 9, 16, ic.MITOS_FP32)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_c_caller_builds(tmp_path):
    """the reference-style C caller compiles against include/mitos_hip.h and
    the replacement include/fft_kernels.h and links against the library"""
    import __graft_entry__ as g
    out = g.build_c_caller(str(tmp_path / "dropin_caller"))
    assert os.path.exists(out)


def test_replacement_fft_kernels_header_is_exported(L):
    """every callback the replacement include/fft_kernels.h declares is an
    exported symbol of the library (so reference units reach the device)"""
    src = open(os.path.join(ROOT, "include", "fft_kernels.h")).read()
    names = re.findall(r"^void (kernel_\w+)\(", src, flags=re.M)
    assert len(names) >= 16
    assert all(hasattr(L, n) for n in names), [n for n in names if not hasattr(L, n)]
