"""Pins of the CPU oracle (CPU only, no GPU).

Each test checks one oracle function against something that does not come
from the oracle itself: the reference's own RNG header (golden draws), an
independent Python/numpy restatement, a closed form, a library routine
(numpy pocketfft, scipy quadrature) or brute force on a tiny input.
"""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def xo():
    with np.load(os.path.join(GOLDEN, "xoshiro.npz")) as f:
        return {k: f[k] for k in f.files}


# ---- a1: xoshiro256** / splitmix64 stream (include/random_xorshift.h:7-59) ------------
def test_draws_match_reference_golden(oracle, xo):
    for seed in (42, 100, 7, 8):
        assert np.array_equal(oracle.draws(seed, 0, 4096), xo[f"first{seed}"]), seed
    for s in range(1, 97):
        assert np.array_equal(oracle.draws(s, 0, 16), xo["seeds1to96"][s - 1]), s


def test_draws_at_offsets_match_reference_golden(oracle, xo):
    for skip, ref in zip(xo["off_skips"], xo["off_draws"]):
        assert np.array_equal(oracle.draws(42, int(skip), 8), ref), int(skip)


def test_state_after_continues_stream(oracle, xo):
    # the state after n draws, stepped once more, gives draw n
    st = oracle.state_after(42, 10)
    s = [int(v) for v in st]
    M = (1 << 64) - 1
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M
    result = (rotl((s[1] * 5) & M, 7) * 9) & M
    assert result == int(xo["first42"][10])


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref",
                                                    "xoshiro_ref")),
                    reason="oracle/_ref is built only where the reference tree exists")
def test_reference_build_reproduces_golden(oracle, xo):
    assert np.array_equal(oracle.ref_draws(42, 0, 4096), xo["first42"])
    assert np.array_equal(oracle.ref_draws(100, int(xo["off_skips"][3]), 8),
                          oracle.draws(100, int(xo["off_skips"][3]), 8))


# ---- a2: Box-Muller (src/random.c:34-49) ----------------------------------------------
def test_box_muller_matches_python_restatement(oracle):
    with np.load(os.path.join(GOLDEN, "normals.npz")) as f:
        ref = f["seed42"]
    z = oracle.normals(42, 0, ref.size)
    assert np.array_equal(z.view(np.uint64), ref.view(np.uint64))


def test_box_muller_uniforms_are_exact():
    # (double)A + 0.5 rounds A to 53 bits first; RMax = 2^64 exactly
    rmax = float(2 ** 64 - 1) + 1.0
    assert rmax == 2.0 ** 64
    assert (float(2 ** 64 - 1) + 0.5) / rmax == 1.0  # largest draw maps to u = 1
    assert (float(0) + 0.5) / rmax == 2.0 ** -65


# ---- a3-a5: GRF draw order and Hermitian planes (src/grf.c:26-175) ----------------------
def _grf_py(N, L, draws):
    """src/grf.c:33-63 restated: x, y, z <= N/2 loop, DC without draws, 4 draws/mode."""
    f = math.sqrt(L ** 3 / 2)
    rmax = 2.0 ** 64
    nrm = lambda a, b: math.sqrt(-2 * math.log((float(a) + 0.5) / rmax)) * \
        math.cos(2 * math.pi * (float(b) + 0.5) / rmax)
    out = np.zeros((N, N, N // 2 + 1), np.complex128)
    j = 0
    for x in range(N):
        for y in range(N):
            for z in range(N // 2 + 1):
                if x == 0 and y == 0 and z == 0:
                    continue
                d = draws[4 * j:4 * j + 4]
                out[x, y, z] = complex(nrm(d[0], d[1]) * f, nrm(d[2], d[3]) * f)
                j += 1
    return out


def test_grf_mode_order(oracle, xo):
    N, L = 8, 100.0
    g = oracle.grf(N, L, 42, hermitian=False)
    ref = _grf_py(N, L, xo["first42"])
    assert g[0, 0, 0] == 0
    assert np.array_equal(g, ref)


def test_grf_hermitian_planes(oracle):
    N, L = 16, 64.0
    raw = oracle.grf(N, L, 5, hermitian=False)
    g = oracle.grf(N, L, 5, hermitian=True)
    neg = lambda i: (-i) % N
    for z in (0, N // 2):
        for x in range(N):
            for y in range(N):
                assert g[x, y, z] == np.conj(g[neg(x), neg(y), z])
                if x == neg(x) and y == neg(y):
                    assert g[x, y, z].imag == 0
    # values with x > N/2 are kept; only the lower half and the x = 0, N/2 lines change
    assert np.array_equal(g[N // 2 + 1:, :, :], raw[N // 2 + 1:, :, :])
    assert np.array_equal(g[:, :, 1:N // 2], raw[:, :, 1:N // 2])


# ---- a8/a9: FFTs (src/fft.c:219-254), against numpy pocketfft and brute force ----------
@pytest.mark.parametrize("N", [8, 12, 16, 32])
def test_fft_r2c_vs_numpy(oracle, N):
    rng = np.random.default_rng(N)
    box = rng.standard_normal((N, N, N))
    ref = np.fft.rfftn(box)
    out = oracle.fft_r2c(box)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def test_fft_r2c_brute_force(oracle):
    N = 4
    rng = np.random.default_rng(1)
    box = rng.standard_normal((N, N, N))
    out = oracle.fft_r2c(box)
    idx = np.arange(N)
    for kx in range(N):
        for ky in range(N):
            for kz in range(N // 2 + 1):
                ph = np.exp(-2j * np.pi * (np.add.outer(np.add.outer(kx * idx, ky * idx),
                                                        kz * idx)) / N)
                assert abs(out[kx, ky, kz] - (box * ph).sum()) < 1e-12


@pytest.mark.parametrize("N", [8, 16, 32])
def test_fft_c2r_vs_numpy_including_nyquist_planes(oracle, N):
    rng = np.random.default_rng(N + 1)
    # a non-Hermitian spectrum (as after i k at the Nyquist index): c2c over x, y,
    # then a c2r along z that drops Im at z' = 0 and z' = N/2
    fbox = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    ref = np.fft.irfftn(fbox, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    out = oracle.fft_c2r(fbox)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def test_fft_plane_wave_closed_form(oracle):
    N = 16
    a, b, c = 3, 5, 2
    i = np.arange(N)
    box = np.cos(2 * np.pi * (a * i[:, None, None] + b * i[None, :, None] + c * i[None, None, :]) / N)
    out = oracle.fft_r2c(box)
    exp = np.zeros_like(out)
    exp[a, b, c] = N ** 3 / 2
    assert np.abs(out - exp).max() < 1e-10


def test_fft_round_trip_normalised(oracle):
    # fft_r2c_dg then fft_c2r_dg returns the input (tests/test_fft.c:60-72)
    N, L = 16, 50.0
    g = oracle.Grid(N, L)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((N, N, N))
    g.box[:, :, :N] = x
    assert oracle.lib().orc_fft_r2c_dg(g.ref()) == 0
    assert g.momentum_space == 1
    assert np.abs(g.fbox - np.fft.rfftn(x) * L ** 3 / N ** 3).max() < 1e-10
    assert oracle.lib().orc_fft_c2r_dg(g.ref()) == 0
    assert np.abs(g.box[:, :, :N] - x).max() < 1e-12


# ---- a10-a16: k-space kernels, closed forms (include/fft_kernels.h) --------------------
def _kgrid(N, L):
    dk = 2 * np.pi / L
    f = np.arange(N)
    f = np.where(f > N // 2, f - N, f)  # index N/2 -> +N/2 (src/fft.c:31-33)
    kx = (f * dk)[:, None, None]
    ky = (f * dk)[None, :, None]
    kz = (np.arange(N // 2 + 1) * dk)[None, None, :]
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    return kx, ky, kz, k


def test_kernel_values(oracle):
    N, L = 16, 80.0
    rng = np.random.default_rng(4)
    fb = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    kx, ky, kz, k = _kgrid(N, L)
    ks = np.where(k > 0, k, 1.0)
    cases = [
        (oracle.K_CONSTANT, (2.5,), 2.5 + 0 * k),
        (oracle.K_INV_POISSON, (), np.where(k > 0, -1.0 / ks / ks, 1.0)),
        (oracle.K_DX, (), 1j * kx + 0 * k), (oracle.K_DY, (), 1j * ky + 0 * k),
        (oracle.K_DZ, (), 1j * kz + 0 * k),
        (oracle.K_SQRT_POWER_LAW, (3.0, 0.965), np.where(k > 0, np.sqrt(3.0 * ks ** 0.965), 0)),
    ]
    for kid, prm, kern in cases:
        out = oracle.apply_kernel(fb, L, kid, prm)
        assert np.abs(out - fb * kern).max() <= 1e-13 * np.abs(fb * kern).max(), kid
    # TSC/CIC window: 1/[sinc(kx L/2N) sinc(ky L/2N) sinc(kz L/2N)]^order, 1 at k = 0
    sinc = lambda x: np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1, x))
    for order in (2, 3):
        w = (sinc(kx * L / (2 * N)) * sinc(ky * L / (2 * N)) * sinc(kz * L / (2 * N))) ** order
        out = oracle.apply_kernel(fb, L, oracle.K_UNDO_WINDOW, (order, N, L))
        assert np.abs(out - fb / w).max() <= 1e-12 * np.abs(fb / w).max()


def test_poisson_inversion(oracle):
    # sum_i d_i d_i (inverse Poisson delta) = delta for k != 0 (tests/test_poisson.c:182-190)
    N, L = 16, 60.0
    rng = np.random.default_rng(5)
    d = np.fft.rfftn(rng.standard_normal((N, N, N)))
    d[0, 0, 0] = 0
    phi = oracle.apply_kernel(d, L, oracle.K_INV_POISSON)
    lap = 0
    for kid in (oracle.K_DX, oracle.K_DY, oracle.K_DZ):
        lap = lap + oracle.apply_kernel(oracle.apply_kernel(phi, L, kid), L, kid)
    assert np.abs(lap - d).max() <= 1e-12 * np.abs(d).max()


def test_spectral_derivative_vs_analytic(oracle):
    # a band-limited field: the spectral derivative is exact
    N, L = 32, 10.0
    i = np.arange(N) * L / N
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    q = 2 * np.pi / L
    f = np.sin(3 * q * X) * np.cos(2 * q * Y) + np.cos(q * Z)
    dfdx = 3 * q * np.cos(3 * q * X) * np.cos(2 * q * Y)
    out = np.fft.irfftn(oracle.apply_kernel(np.fft.rfftn(f), L, oracle.K_DX), s=(N, N, N),
                       axes=(0, 1, 2))
    assert np.abs(out - dfdx).max() < 1e-11


def test_2lpt_two_orthogonal_modes(oracle):
    # delta = a cos(k1 x) + b cos(k2 y): phi1 = -delta/k^2, delta2 = phi_xx phi_yy
    # = a b cos(k1 x) cos(k2 y), phi2 = -delta2/(k1^2 + k2^2); pot = f1 phi1 + f2 phi2
    N, L, a, b, m1, m2, f1, f2 = 16, 40.0, 0.7, -0.4, 2, 3, 1.0, -3.0 / 7.0
    q = 2 * np.pi / L
    k1, k2 = m1 * q, m2 * q
    i = np.arange(N) * L / N
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    delta = a * np.cos(k1 * X) + b * np.cos(k2 * Y)
    dens, pot, work = oracle.Grid(N, L), oracle.Grid(N, L), oracle.Grid(N, L)
    dens.box[:, :, :N] = delta
    assert oracle.lib().orc_fft_r2c_dg(dens.ref()) == 0
    assert oracle.solve_2lpt(pot, dens, work, f1, f2) == 0
    assert pot.momentum_space == 1
    assert oracle.lib().orc_fft_c2r_dg(pot.ref()) == 0
    phi1 = -a / k1 ** 2 * np.cos(k1 * X) - b / k2 ** 2 * np.cos(k2 * Y)
    phi2 = -a * b * np.cos(k1 * X) * np.cos(k2 * Y) / (k1 ** 2 + k2 ** 2)
    exp = f1 * phi1 + f2 * phi2
    assert np.abs(pot.box[:, :, :N] - exp).max() < 1e-12 * np.abs(exp).max()


def test_2lpt_single_mode_has_no_second_order_term(oracle):
    # one Fourier mode has a rank-one Hessian, so delta2 = 0 identically
    N, L = 16, 40.0
    q = 2 * np.pi / L
    i = np.arange(N) * L / N
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    delta = np.cos(q * (X + 2 * Y - Z))
    out = []
    for f2 in (0.0, -3.0 / 7.0):
        dens, pot, work = oracle.Grid(N, L), oracle.Grid(N, L), oracle.Grid(N, L)
        dens.box[:, :, :N] = delta
        oracle.lib().orc_fft_r2c_dg(dens.ref())
        oracle.solve_2lpt(pot, dens, work, 1.0, f2)
        out.append(pot.fbox.copy())
    assert np.abs(out[0] - out[1]).max() < 1e-12 * np.abs(out[0]).max()


# ---- a11: BBKS and the sigma8 amplitude (scipy quadrature) ------------------------------
def test_bbks_closed_form(oracle):
    for k in (1e-3, 0.05, 0.3, 2.0):
        q = k / 0.21
        T = math.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3
                                                     + (6.71 * q) ** 4) ** -0.25
        assert abs(oracle.bbks(k, 0.21) - T) < 1e-15 * T


def test_sigma8_amplitude_vs_quadrature(oracle):
    from scipy import integrate

    n_s, gamma, s8, R = 0.965, 0.21, 0.81, 8.0

    def T(k):
        q = k / gamma
        return math.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 +
                                                       (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25

    def W(x):
        return 3 * (math.sin(x) - x * math.cos(x)) / x ** 3 if x > 1e-3 else 1 - x * x / 10

    f = lambda lnk: (math.exp(lnk) ** (3 + n_s) * T(math.exp(lnk)) ** 2 *
                     W(math.exp(lnk) * R) ** 2 / (2 * math.pi ** 2))
    # adaptive quadrature on 300 pieces of ln k (the window oscillates for kR >> 1)
    edges = np.linspace(math.log(1e-6), math.log(1e3), 301)
    with np.errstate(all="ignore"):
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            I = sum(integrate.quad(f, a, b, epsabs=0, epsrel=1e-12, limit=200)[0]
                    for a, b in zip(edges[:-1], edges[1:]))
    A = s8 ** 2 / I
    assert abs(oracle.sigma8_amplitude(n_s, gamma, s8, R) / A - 1) < 1e-10


# ---- a17-a19: lattice, TSC gather, CIC ---------------------------------------------------
def test_lattice_closed_form(oracle):
    M, L = 6, 100.0
    X, Y, Z, ids = oracle.lattice(M, L, 1)
    sp = np.float32(L) / np.float32(M)  # float spacing (src/particle.c:102-103)
    q = np.arange(M ** 3)
    x, y, z = q // (M * M), (q // M) % M, q % M
    assert np.array_equal(X, (x * sp).astype(np.float32))
    assert np.array_equal(Y, (y * sp).astype(np.float32))
    assert np.array_equal(Z, (z * sp).astype(np.float32))
    assert np.array_equal(ids, q + 1)


def _padded(f):
    N = f.shape[0]
    b = np.zeros((N, N, N + 2))
    b[:, :, :N] = f
    return b


def test_tsc_constant_and_linear_fields(oracle):
    N, L = 16, 32.0
    c = _padded(np.full((N, N, N), 2.5))
    rng = np.random.default_rng(6)
    for _ in range(20):
        x, y, z = rng.uniform(0, L, 3)
        assert abs(oracle.grid_tsc(c, L, x, y, z) - 2.5) < 1e-14
    # TSC reproduces a linear field away from the periodic seam
    i = np.arange(N)
    lin = _padded(np.broadcast_to(0.3 * i[:, None, None] - 0.2 * i[None, :, None] +
                                  0.1 * i[None, None, :], (N, N, N)).copy())
    for _ in range(20):
        x, y, z = rng.uniform(4, 10, 3)
        X, Y, Z = x * N / L, y * N / L, z * N / L
        assert abs(oracle.grid_tsc(lin, L, x, y, z) - (0.3 * X - 0.2 * Y + 0.1 * Z)) < 1e-12


def test_tsc_weights_at_a_node_and_midpoint(oracle):
    N, L = 8, 8.0
    f = np.zeros((N, N, N))
    f[3, 3, 3] = 1.0
    b = _padded(f)
    assert abs(oracle.grid_tsc(b, L, 3.0, 3.0, 3.0) - 0.75 ** 3) < 1e-15
    assert abs(oracle.grid_tsc(b, L, 4.0, 3.0, 3.0) - 0.125 * 0.75 ** 2) < 1e-15
    assert abs(oracle.grid_tsc(b, L, 3.5, 3.0, 3.0) - 0.5 * 0.75 ** 2) < 1e-15


def test_displacement_is_sequential(oracle):
    # psi_x = const; psi_y varies with x only: the y move reads psi_y at the moved x
    N, L = 16, 16.0
    i = np.arange(N)
    px = np.full((N, N, N), 1.0)
    py = np.broadcast_to(0.5 * i[:, None, None], (N, N, N)).copy()
    pz = np.zeros((N, N, N))
    psi = np.stack([_padded(px), _padded(py), _padded(pz)])
    X = np.array([6.0], np.float32)
    Y = np.array([7.0], np.float32)
    Z = np.array([2.0], np.float32)
    oracle.displace_tsc(psi, L, X, Y, Z)
    assert X[0] == np.float32(5.0)
    assert abs(Y[0] - (7.0 - 0.5 * 5.0)) < 1e-6  # psi_y at x = 5, not at x = 6
    assert Z[0] == np.float32(2.0)


def test_cic_mass_conservation_and_weights(oracle):
    N, L = 8, 16.0
    box, tot = oracle.cic(N, L, np.array([3.5], np.float32), np.array([4.0], np.float32),
                          np.array([9.25], np.float32), mass=1.0)
    cellvol = (L / N) ** 3
    # X = 1.75, Y = 2, Z = 4.625 cells
    exp = np.zeros((N, N, N))
    for dx, wx in ((1, 0.25), (2, 0.75)):
        for dz, wz in ((4, 0.375), (5, 0.625)):
            exp[dx, 2, dz] += wx * wz / cellvol
    assert np.abs(box - exp).max() < 1e-15
    rng = np.random.default_rng(7)
    n = 1000
    X, Y, Z = (rng.uniform(0, L, n).astype(np.float32) for _ in range(3))
    box, tot = oracle.cic(N, L, X, Y, Z, mass=1.0)
    assert abs(box.sum() * cellvol - n) < 1e-9


def test_cic_unperturbed_lattice_is_uniform(oracle):
    N, L = 8, 20.0
    X, Y, Z, _ = oracle.lattice(N, L, 1)
    box, tot = oracle.cic(N, L, X, Y, Z, mass=1.0)
    oracle.overdensity(box, L, tot)
    assert np.abs(box).max() < 1e-12


def test_pipeline_particle_lattice_m(oracle):
    # CubeRootNumber M (src/particle.c:87-129) in the oracle's pipeline:
    # M = N is the default run bit for bit; an M = 2N lattice barely moved
    # (sigma8 = 1e-9) deposits a uniform density on the N grid, which fixes
    # the lattice spacing L / M and the normalisation by the M^3 particles
    # (an N^3 normalisation would give delta = 7 everywhere); an M = N / 2
    # lattice conserves mass
    N, L = 16, 64.0
    a = oracle.ic_run(N, L, 5, bins=8, want_density=True)
    b = oracle.ic_run(N, L, 5, bins=8, want_density=True, M=N)
    for key in ("P", "X", "Y", "Z", "density"):
        assert np.array_equal(a[key], b[key])
    c = oracle.ic_run(N, L, 5, bins=8, want_density=True, sigma8=1e-9, M=2 * N)
    assert c["X"].size == (2 * N) ** 3
    assert np.abs(c["density"]).max() < 1e-6
    d = oracle.ic_run(N, L, 5, bins=8, want_density=True, M=N // 2)
    assert d["X"].size == (N // 2) ** 3
    assert abs(d["density"].mean()) < 1e-12
    assert d["density"].min() == -1.0  # cells no particle reaches


# ---- a20: P(k) binning, counts and Parseval (src/calc_powerspec.c:32-90) ----------------
def _pk_bins_numpy(N, L, bins):
    """The reference's bin formula restated with numpy (u and (bins-1)*u in float32)."""
    kx, ky, kz, k = _kgrid(N, L)
    dk = 2 * np.pi / L
    lmin, lmax = math.log(dk), math.log(math.sqrt(3) * dk * N / 2)
    with np.errstate(divide="ignore"):
        u = ((np.log(k) - lmin) / (lmax - lmin)).astype(np.float32)
    b = np.floor(np.float32(bins - 1) * u).astype(np.int64)
    b[0, 0, 0] = -1
    return b, k


@pytest.mark.parametrize("N,L,bins", [(16, 100.0, 8), (32, 256.0, 16)])
def test_powerspec_counts_and_parseval(oracle, N, L, bins):
    rng = np.random.default_rng(N)
    d = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    kk, P, obs = oracle.powerspec(d, d, L, bins)
    b, k = _pk_bins_numpy(N, L, bins)
    m = np.full(k.shape, 2)
    m[:, :, 0] = 1
    m[:, :, N // 2] = 1
    for i in range(bins):
        sel = b == i
        assert obs[i] == m[sel].sum(), i
        if obs[i]:
            assert abs(kk[i] - (m[sel] * k[sel]).sum() / obs[i]) < 1e-12 * kk[i]
    ok = obs > 0
    lhs = (P[ok] * obs[ok]).sum() * L ** 3
    rhs = (m * np.abs(d) ** 2)[b >= 0].sum()
    assert abs(lhs / rhs - 1) < 1e-12


def test_powerspec_white_noise_expectation(oracle):
    # E|delta_k|^2 = sigma^2 N^3 for unit white noise (tests/test_calc_powerspec.c:74-98)
    N, L, bins = 32, 32.0, 8
    rng = np.random.default_rng(9)
    d = np.fft.rfftn(rng.standard_normal((N, N, N)))
    kk, P, obs = oracle.powerspec(d, d, L, bins)
    ok = obs > 500
    mean = P[ok] * L ** 3 / N ** 3
    assert np.all(np.abs(mean - 1) < 5 / np.sqrt(obs[ok]) + 0.02)


# ---- the FP64 reference gradient (oracle/adjoint_ref.py) ------------------------------
def test_adjoint_reference_vs_finite_differences(oracle):
    import torch
    from oracle import adjoint_ref as R
    N, L, bins = 8, 64.0, 6
    fw = R.Forward(N, L, bins, oracle.sigma8_amplitude(0.965, 0.21, 0.81))
    rng = np.random.default_rng(1)
    H = N // 2 + 1
    noise = rng.standard_normal((N, N, H)) + 1j * rng.standard_normal((N, N, H))
    with torch.no_grad():
        P0 = fw.power(fw.density(torch.tensor(noise))).numpy()
    tgt = P0 * (1 + 0.1 * rng.standard_normal(bins))
    _, g = R.pk_gradient(fw, noise, tgt)

    def loss(nz):
        with torch.no_grad():
            return float(R.pk_loss(fw.power(fw.density(torch.tensor(nz))), tgt))

    for (x, y, z) in [(1, 2, 1), (3, 5, 2), (6, 1, 3), (5, 5, 0), (2, 2, 0), (0, 1, 4)]:
        for d in (1e-6, 1e-6j):
            p = noise.copy()
            p[x, y, z] += d
            lp = loss(p)
            p[x, y, z] -= 2 * d
            fd = (lp - loss(p)) / 2e-6
            ga = g[x, y, z].real if d.real else g[x, y, z].imag
            assert abs(fd - ga) <= 1e-6 * np.abs(g).max(), (x, y, z, d, fd, ga)


def test_adjoint_reference_forward_matches_oracle_pipeline(oracle):
    # the torch restatement of the forward (stencil form, float64 positions)
    # agrees with the oracle's reference-order pipeline to the rounding of the
    # oracle's float positions
    import torch
    from oracle import adjoint_ref as R
    N, L, bins = 16, 64.0, 8
    A = oracle.sigma8_amplitude(0.965, 0.21, 0.81)
    fw = R.Forward(N, L, bins, A)
    # the seed-42 noise: the oracle GRF with sqrt(L^3/2) = 1
    noise = oracle.grf(N, 2.0 ** (1.0 / 3.0), 42, hermitian=False)
    with torch.no_grad():
        P = fw.power(fw.density(torch.tensor(noise))).numpy()
    ref = oracle.ic_run(N, L, 42, bins=bins)
    ok = ref["obs"] > 0
    assert np.array_equal(fw.obs[ok], ref["obs"][ok])
    assert np.max(np.abs(P[ok] / ref["P"][ok] - 1)) < 1e-4


# ---- a11: kernel_power_no_transfer (include/fft_kernels.h:116-121) ---------------------
def test_power_no_transfer_invariants(oracle):
    """kern = sqrt(P) with P = k^4 P_phi the bare primordial spectrum of
    src/primordial.c:25-33: P(k_pivot) = 2 pi^2 A_s k_pivot, the log-slope of
    P is n_s (P = A_s k_p^(1-n_s) 2 pi^2 k^n_s), P(0) = 0 -- checked from
    the oracle's values, without restating its expression"""
    A_s, n_s, kp = 2.1e-9, 0.965, 0.05
    N = 16
    # boxes whose fundamental mode is k_pivot and 3 k_pivot
    for L, scale in ((2 * math.pi / kp, 1.0), (2 * math.pi / (3 * kp), 3.0)):
        fb = np.ones((N, N, N // 2 + 1), dtype=np.complex128)
        out = oracle.apply_kernel(fb, L, oracle.K_POWER_NO_TRANSFER, (A_s, n_s, kp))
        assert out[0, 0, 0] == 0
        P1 = out[0, 0, 1].real ** 2  # k = k_fund
        P2 = out[0, 0, 2].real ** 2  # k = 2 k_fund
        assert np.all(out.imag == 0)
        if scale == 1.0:
            assert abs(P1 / (2 * math.pi ** 2 * A_s * kp) - 1) < 1e-14
        assert abs(math.log(P2 / P1) / math.log(2.0) - n_s) < 1e-13
        # isotropy: the same |k| along x, y and z gives the same value
        assert out[1, 0, 0] == out[0, 1, 0] == out[0, 0, 1]
    # between boxes: P(3 k_p) / P(k_p) = 3^n_s
    Lb = 2 * math.pi / kp
    fb = np.ones((N, N, N // 2 + 1), dtype=np.complex128)
    out = oracle.apply_kernel(fb, Lb, oracle.K_POWER_NO_TRANSFER, (A_s, n_s, kp))
    assert abs(out[0, 0, 3].real ** 2 / out[0, 0, 1].real ** 2 - 3 ** n_s) < 1e-13


# ---- end to end: the realised spectrum recovers the input -----------------------------
def _modes_per_bin(N, L, bins, Pfun):
    """mean over each bin's modes (multiplicity 1 on z = 0, N/2, else 2) of
    Pfun(k), with the bins of the reference's formula"""
    b, k = _pk_bins_numpy(N, L, bins)
    m = np.full(k.shape, 2.0)
    m[:, :, 0] = 1
    m[:, :, N // 2] = 1
    Pk = np.where(k > 0, Pfun(np.where(k > 0, k, 1.0)), 0.0)
    exp = np.zeros(bins)
    cnt = np.zeros(bins)
    for i in range(bins):
        sel = b == i
        exp[i] = (m[sel] * Pk[sel]).sum()
        cnt[i] = m[sel].sum()
    return np.where(cnt > 0, exp / np.maximum(cnt, 1), np.nan), cnt


def test_grf_recovers_input_spectrum(oracle):
    """tests/test_calc_powerspec.c:47-98: GRF, multiplied by sqrt(P(k)),
    through c2r and r2c with the reference's normalisations, binned: every bin
    with enough modes recovers P within cosmic variance (independent complex
    modes: relative scatter sqrt(2/obs)); the chi^2 over the bins is ~1 per
    bin.  This pins the GRF amplitude sqrt(L^3/2), the c2r division by L^3,
    the r2c factor L^3/N^3 and the P(k) division by L^3 together."""
    N, L, bins = 64, 400.0, 24
    A, ns = 2.0e4, 0.97
    f = oracle.grf(N, L, 17)
    f = oracle.apply_kernel(f, L, oracle.K_SQRT_POWER_LAW, (A, ns))
    box = oracle.fft_c2r(f) / L ** 3                     # src/fft.c:110-125
    g = oracle.fft_r2c(box) * (L ** 3 / N ** 3)           # src/fft.c:38-49
    kk, P, obs = oracle.powerspec(g, g, L, bins)
    exp, cnt = _modes_per_bin(N, L, bins, lambda k: A * k ** ns)
    ok = obs > 40
    assert np.array_equal(obs, cnt.astype(obs.dtype))
    r = P[ok] / exp[ok] - 1
    assert np.all(np.abs(r) < 5 * np.sqrt(2.0 / obs[ok]))
    chi2 = float(np.sum(r ** 2 * obs[ok] / 2) / ok.sum())
    assert 0.3 < chi2 < 2.5, chi2


def test_pipeline_low_k_recovers_input_spectrum(oracle):
    """the oracle's whole FP64 pipeline (GRF, BBKS spectrum normalised to
    sigma8, 2LPT, TSC-deconvolved displacements, lattice particles, CIC,
    P(k) with the CIC window undone): at k well below the non-linear scale
    the particle P(k) follows A k^n_s T_BBKS(k)^2 within cosmic variance"""
    N, L, bins = 64, 2000.0, 24
    res = oracle.ic_run(N, L, 42, bins=bins)
    A = oracle.sigma8_amplitude(0.965, 0.21, 0.81)
    exp, cnt = _modes_per_bin(N, L, bins,
                              lambda k: A * k ** 0.965 * np.vectorize(oracle.bbks)(k, 0.21) ** 2)
    ok = (res["obs"] > 20) & (res["k"] < 0.05)
    assert ok.sum() >= 4
    r = res["P"][ok] / exp[ok] - 1
    assert np.all(np.abs(r) < 5 * np.sqrt(2.0 / res["obs"][ok]) + 0.02)


# ---- a12: tabulated transfer functions (src/perturb_spline.c) -------------------------
def _synthetic_table(rng, nk=57, nt=6, nf=3, kmin=0.004, kmax=3.0):
    k = np.sort(np.exp(rng.uniform(np.log(kmin), np.log(kmax), nk)))
    k[0], k[-1] = kmin, kmax
    lt = np.sort(rng.uniform(-2.0, 3.0, nt))
    delta = rng.standard_normal((nf, nt, nk)) + 3.0
    return k, lt, delta


def _chord(k, row, kk):
    """brute force: the interval with k_i < kk <= k_{i+1} by searchsorted
    (first interval at k_0), clamped to the last one above the table"""
    if kk > k[-1]:
        return row[-1]
    i = max(int(np.searchsorted(k, kk, side="left")) - 1, 0)
    u = (kk - k[i]) / (k[i + 1] - k[i])
    return (1 - u) * row[i] + u * row[i + 1]


def test_transfer_search_and_interp_brute_force(oracle):
    rng = np.random.default_rng(21)
    k, lt, delta = _synthetic_table(rng)
    T = oracle.Transfer(k, lt, delta, acc_size=37)
    # nodes: the table value itself, from either neighbouring interval
    for j in (0, 1, 17, len(k) - 2, len(k) - 1):
        ki, uk = T.find_k(k[j])
        assert T.interp(ki, 2, uk, 0.0, 1) == delta[1, 2, j]
    # random k and tau against the brute-force chord in k and the linear rule in tau
    for kk in np.exp(rng.uniform(np.log(0.004), np.log(5.0), 400)):
        ki, uk = T.find_k(kk)
        for ltv in (lt[0] - 1, lt[1] + 0.3 * (lt[2] - lt[1]), lt[-1] + 2):
            ti, ut = T.find_tau(ltv)
            got = T.interp(ki, ti, uk, ut, 2)
            lo = _chord(k, delta[2, ti], kk)
            hi = _chord(k, delta[2, ti + 1], kk)
            ref = (1 - ut) * lo + ut * hi
            assert abs(got - ref) <= 1e-13 * abs(ref), (kk, ltv)
    # tau outside the table: the first row / the last row
    assert T.find_tau(lt[0] - 5) == (0, 0.0)
    assert T.find_tau(lt[-1] + 5) == (len(lt) - 2, 1.0)
    assert T.find_tau(lt[-1]) == (len(lt) - 2, 1.0)


def test_transfer_kernel_on_grid(oracle):
    """kernel_transfer_function applied to a half grid: DC -> 0, every other
    mode multiplied by the interpolated T(|k|); the inverse kernel by 1/T"""
    rng = np.random.default_rng(22)
    k, lt, delta = _synthetic_table(rng, kmin=0.01, kmax=0.9)
    T = oracle.Transfer(k, lt, delta)
    N, L = 16, 120.0
    fb = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    ti, ut = T.find_tau(0.4)
    out = T.apply(fb, L, 1, ti, ut)
    inv = T.apply(fb, L, 1, ti, ut, invert=True)
    _, _, _, kk = _kgrid(N, L)
    assert out[0, 0, 0] == 0 and inv[0, 0, 0] == 0
    for idx in [(0, 0, 1), (3, 5, 2), (N // 2, N // 2, N // 2), (N - 1, 1, 7)]:
        kv = kk[idx]
        t = (1 - ut) * _chord(k, delta[1, ti], kv) + ut * _chord(k, delta[1, ti + 1], kv)
        assert abs(out[idx] - fb[idx] * t) <= 1e-13 * abs(fb[idx] * t)
        assert abs(inv[idx] - fb[idx] / t) <= 1e-13 * abs(fb[idx] / t)


# ---- f4: shrinkGrid (src/shrink_grids.c:55-75) ------------------------------------------
def test_shrink_grid_block_means(oracle):
    rng = np.random.default_rng(23)
    N, M = 12, 4
    box = rng.standard_normal((N, N, N))
    out = oracle.shrink_grid(box, M)
    K = N // M
    ref = box.reshape(M, K, M, K, M, K).mean(axis=(1, 3, 5))
    assert np.abs(out - ref).max() < 1e-14
    assert abs(out.sum() * K ** 3 - box.sum()) < 1e-11


# ---- the remaining kernels of include/fft_kernels.h:33-93 -------------------------------
def test_filter_kernels_invariants(oracle):
    N, L = 16, 100.0
    dk = 2 * math.pi / L
    fb = np.ones((N, N, N // 2 + 1), dtype=np.complex128)
    _, _, _, k = _kgrid(N, L)
    n2 = np.rint((k / dk) ** 2).astype(np.int64)  # integer |n|^2 of every mode
    lp = oracle.apply_kernel(fb, L, oracle.K_LOWPASS, (math.sqrt(20.5) * dk,))
    hp = oracle.apply_kernel(fb, L, oracle.K_HIPASS, (math.sqrt(20.5) * dk,))
    th = oracle.apply_kernel(fb, L, oracle.K_TOPHAT, (math.sqrt(5.5) * dk, math.sqrt(20.5) * dk))
    assert np.array_equal(lp.real, (n2 < 20.5).astype(float))
    assert np.array_equal(hp.real, (n2 > 20.5).astype(float))
    assert np.array_equal(th.real, ((n2 > 5.5) & (n2 < 20.5)).astype(float))
    # an elliptic top hat with eta2 = 1 is the sphere |k| |r| <= 2 pi
    r = (3.0, 4.0, 12.0)  # |r| = 13
    et = oracle.apply_kernel(fb, L, oracle.K_ELLIPTIC_TOPHAT, r + (1.0,))
    assert np.array_equal(et.real, (k * 13.0 <= 2 * math.pi + 1e-12).astype(float))
    # Gaussian and its inverse: product 1, value 1/e at k R = 1
    R = 7.0
    g = oracle.apply_kernel(fb, L, oracle.K_GAUSSIAN, (R,))
    gi = oracle.apply_kernel(fb, L, oracle.K_GAUSSIAN_INV, (R,))
    assert np.abs(g * gi - 1).max() < 1e-12
    assert abs(g[0, 0, 0] - 1) == 0
    # translation: unit modulus, composes additively, r = L is the identity
    t1 = oracle.apply_kernel(fb, L, oracle.K_TRANSLATE, (1.5, -2.0, 0.25))
    t2 = oracle.apply_kernel(t1, L, oracle.K_TRANSLATE, (0.5, 2.0, 0.75))
    t12 = oracle.apply_kernel(fb, L, oracle.K_TRANSLATE, (2.0, 0.0, 1.0))
    assert np.abs(np.abs(t1) - 1).max() < 1e-14
    assert np.abs(t2 - t12).max() < 1e-12
    tL = oracle.apply_kernel(fb, L, oracle.K_TRANSLATE, (L, 0.0, 0.0))
    assert np.abs(tL - 1).max() < 1e-12


# ---- f3: Monge-Ampere (src/monge_ampere.c:50-183) --------------------------------------
def _ma_grids(oracle, N, L, delta):
    dens, pot, work = oracle.Grid(N, L), oracle.Grid(N, L), oracle.Grid(N, L)
    dens.box[:, :, :N] = delta
    dens.fbox[:] = oracle.fft_r2c(delta) * (L ** 3 / N ** 3)
    dens.s.momentum_space = b"\x01"
    return dens, pot, work


def test_monge_ampere_one_dimensional_is_poisson(oracle):
    """for a density that varies along x only, det(I + D^2 phi) = 1 + phi_xx,
    so the Zel'dovich (Poisson) potential already solves the equation: the
    residual is at rounding level and further cycles leave phi unchanged"""
    N, L = 16, 64.0
    x = np.arange(N) * L / N
    delta = np.broadcast_to((0.3 * np.cos(2 * np.pi * 2 * x / L) +
                             0.2 * np.sin(2 * np.pi * 5 * x / L))[:, None, None],
                            (N, N, N)).copy()
    dens, pot, work = _ma_grids(oracle, N, L, delta)
    rc, eps = oracle.solve_monge_ampere(pot, dens, work, 3)
    assert rc == 0
    assert np.all(eps < 1e-13)
    dens0, pot0, work0 = _ma_grids(oracle, N, L, delta)
    rc, _ = oracle.solve_monge_ampere(pot0, dens0, work0, 0)
    assert np.abs(pot.fbox - pot0.fbox).max() < 1e-12 * np.abs(pot0.fbox).max()


def test_monge_ampere_residual_decreases(oracle):
    """a small three-dimensional perturbation: each cycle lowers the residual
    of det(I + D^2 phi) = 1 + delta (by a factor of about 30 over six cycles)"""
    N, L = 16, 64.0
    rng = np.random.default_rng(12)
    f = oracle.apply_kernel(rng.standard_normal((N, N, N // 2 + 1)) +
                            1j * rng.standard_normal((N, N, N // 2 + 1)), L,
                            oracle.K_GAUSSIAN, (4.0,))
    delta = oracle.fft_c2r(f)
    delta *= 0.2 / np.abs(delta).max()
    dens, pot, work = _ma_grids(oracle, N, L, delta)
    rc, eps = oracle.solve_monge_ampere(pot, dens, work, 6)
    assert rc == 0
    assert np.all(np.diff(eps) < 0), eps
    assert eps[-1] < 0.1 * eps[0]
    assert oracle.solve_monge_ampere(pot, work, dens, 1)[0] == 1  # not in momentum space


# ---- a6: RANLUX ranlxd1 and generate_ngeniclike_grf (src/grf_ngeniclike.c) -------------
def _ranlux48_integer_model(seed, n, p=202):
    """RANLUX restated on 48-bit integers (not doubles): the seed register of
    31 bits with taps 0 and 18 gives x_{-12..-1}; then x_t = x_{t-5} - x_{t-12}
    - c mod 2^48 with borrow c; each refill of p steps delivers its last 12
    values, oldest first"""
    if seed == 0:
        seed = 1
    v = seed & 0xFFFFFFFF
    bits = [(v >> k) & 1 for k in range(31)]
    ib, jb = 0, 18
    seq = []
    for _ in range(12):
        a = 0
        for _ in range(48):
            a = 2 * a + (bits[ib] ^ 1)
            bits[ib] ^= bits[jb]
            ib, jb = (ib + 1) % 31, (jb + 1) % 31
        seq.append(a)
    out, c = [], 0
    while len(out) < n:
        for _ in range(p):
            d = seq[-5] - seq[-12] - c
            c = 1 if d < 0 else 0
            seq.append(d + (1 << 48) if d < 0 else d)
        out.extend(seq[-12:])
        seq = seq[-12:]
    return np.array(out[:n], dtype=np.float64) / 2.0 ** 48


def test_ranlxd_double_stream_matches_integer_model(oracle):
    for seed in (1, 42, 12345, 0x7ffffffe):
        got = oracle.ranlxd_draws(seed, 600)
        assert np.array_equal(got, _ranlux48_integer_model(seed, 600)), seed
    assert np.array_equal(oracle.ranlxd_draws(0, 30), oracle.ranlxd_draws(1, 30))
    u = oracle.ranlxd_draws(7, 200000)
    assert np.all((u >= 0) & (u < 1))
    assert np.all(u * 2.0 ** 48 == np.floor(u * 2.0 ** 48))  # 48-bit fractions
    assert abs(u.mean() - 0.5) < 5 / np.sqrt(12 * u.size)


def test_ngeniclike_grf_structure(oracle):
    """src/grf_ngeniclike.c:116-195: zero DC, Nyquist planes x, y = N/2 and
    the z = N/2 plane; conjugate partners on the z = 0 plane; |delta|^2 /
    L^3 = -ln(ampl) is Exp(1)-distributed, the phases uniform"""
    N, L = 32, 150.0
    f = oracle.ngeniclike_grf(N, L, 42)
    assert f[0, 0, 0] == 0
    assert np.all(f[N // 2] == 0) and np.all(f[:, N // 2] == 0) and np.all(f[:, :, N // 2] == 0)
    z0 = f[:, :, 0]
    m = (-np.arange(N)) % N
    for x in range(N):
        for y in range(N):
            if x in (0, N // 2) or y == N // 2:
                continue
            assert z0[m[x], m[y]] == np.conj(z0[x, y])
    for y in range(1, N // 2):
        assert z0[0, N - y] == np.conj(z0[0, y])
    inner = np.abs(f[1:N // 2, 1:N // 2, 1:N // 2]).ravel() ** 2 / L ** 3
    assert abs(inner.mean() - 1) < 5 / np.sqrt(inner.size)
    assert abs(np.mean(inner > 1) - np.exp(-1)) < 5 * 0.5 / np.sqrt(inner.size)
    ph = np.angle(f[1:N // 2, 1:N // 2, 1:N // 2]).ravel()
    assert abs(np.mean(np.cos(ph))) < 5 / np.sqrt(ph.size)
    # the seed changes the field; the same seed repeats it
    assert np.array_equal(f, oracle.ngeniclike_grf(N, L, 42))
    assert not np.array_equal(f, oracle.ngeniclike_grf(N, L, 43))
