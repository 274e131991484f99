"""GPU parity of the rest of the drop-in surface against the oracle: the
filter kernels, tabulated transfer functions and generatePerturbationGrid,
host-evaluated callbacks, the serial-array FFT helpers and window
corrections, solvePoisson, the float/MPI P(k) variant, sampleNorm and
gridTSC_dg, particle velocities and the export layout, field files and
shrinkGrid.  Every call goes through the C ABI (ctypes)."""
import ctypes
import math
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _fbox(rng, N):
    return rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))


def test_filter_kernels_vs_oracle(gpu, oracle):
    """include/fft_kernels.h:33-93 on the device against the oracle"""
    rng = np.random.default_rng(31)
    N, L = 32, 150.0
    f = _fbox(rng, N)
    dk = 2 * math.pi / L
    cases = [
        ("kernel_lowpass", {"value": 7.5 * dk}, oracle.K_LOWPASS, (7.5 * dk,), 0.0),
        ("kernel_hipass", {"value": 3.2 * dk}, oracle.K_HIPASS, (3.2 * dk,), 0.0),
        ("kernel_tophat", {"values": (2.5 * dk, 9.1 * dk)}, oracle.K_TOPHAT,
         (2.5 * dk, 9.1 * dk), 0.0),
        ("kernel_elliptic_tophat", {"values": (3.0, 1.0, 2.0, 0.7)}, oracle.K_ELLIPTIC_TOPHAT,
         (3.0, 1.0, 2.0, 0.7), 0.0),
        ("kernel_translate", {"values": (4.5, -2.0, 11.0)}, oracle.K_TRANSLATE,
         (4.5, -2.0, 11.0), 1e-14),
        ("kernel_gaussian", {"value": 6.0}, oracle.K_GAUSSIAN, (6.0,), 1e-14),
        ("kernel_gaussian_inv", {"value": 1.5}, oracle.K_GAUSSIAN_INV, (1.5,), 1e-14),
    ]
    r = gpu.Grid(N, L)
    w = gpu.Grid(N, L)
    for name, kw, kid, prm, tol in cases:
        print("case kernel", name, flush=True)
        r.set_fbox(f)
        r.momentum_space = 1
        p = gpu.kernel_params(name, **kw)
        assert gpu.fft_apply_kernel_dg(w, r, name, p) == 0
        ref = oracle.apply_kernel(f, L, kid, prm)
        err = rel(w.fbox(), ref)
        print("  rel err", err, flush=True)
        assert err <= tol, name


def test_host_evaluated_callback(gpu):
    """an unrecognised callback (here a Python one, as a caller's own static
    inline kernel would be) is evaluated on the host for every mode; a copy of
    kernel_dx gives the device kernel's bits, a caller's own kernel the
    product with its values"""
    from mitos_amd import _lib
    N, L = 16, 50.0
    rng = np.random.default_rng(32)
    f = _fbox(rng, N)
    r, a, b = gpu.Grid(N, L), gpu.Grid(N, L), gpu.Grid(N, L)
    r.set_fbox(f)
    r.momentum_space = 1

    @_lib.KERNEL_FN
    def my_dx(k):
        k.contents.kern[0] = 0.0
        k.contents.kern[1] = k.contents.kx

    @_lib.KERNEL_FN
    def my_shell(k):
        inside = 0.3 < k.contents.k < 0.8
        k.contents.kern[0] = 1.0 if inside else 0.0
        k.contents.kern[1] = -0.25 if inside else 0.0

    assert gpu.fft_apply_kernel_dg(a, r, "kernel_dx") == 0
    code = gpu.lib().fft_apply_kernel_dg(b.ref(), r.ref(), ctypes.cast(my_dx, ctypes.c_void_p),
                                         None)
    assert code == 0 and b.momentum_space == 1
    assert np.array_equal(a.fbox(), b.fbox())
    code = gpu.lib().fft_apply_kernel_dg(b.ref(), r.ref(),
                                         ctypes.cast(my_shell, ctypes.c_void_p), None)
    assert code == 0
    dk = 2 * np.pi / L
    fr = np.fft.fftfreq(N, 1.0 / N)
    fr = np.where(fr == -N // 2, N // 2, fr)
    kx = fr[:, None, None] * dk
    ky = fr[None, :, None] * dk
    kz = (np.arange(N // 2 + 1) * dk)[None, None, :]
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    m = np.where((k > 0.3) & (k < 0.8), 1.0 - 0.25j, 0.0)
    assert np.abs(b.fbox() - f * m).max() < 1e-15 * np.abs(f).max()


def _table(rng, N, L):
    """a synthetic table covering the grid's modes (k_min below the
    fundamental mode) and stopping below its largest |k| (the clamp)"""
    dk = 2 * math.pi / L
    nk, nt = 73, 7
    k = np.sort(np.exp(rng.uniform(np.log(0.5 * dk), np.log(0.7 * math.sqrt(3) * dk * N / 2),
                                   nk)))
    k[0] = 0.5 * dk
    lt = np.sort(rng.uniform(-3.0, 2.0, nt))
    delta = rng.standard_normal((3, nt, nk)) + 4.0
    return k, lt, delta


def test_transfer_kernel_vs_oracle(gpu, oracle):
    rng = np.random.default_rng(33)
    for N, L in ((32, 200.0), (48, 333.0)):
        print("case transfer kernel", N, L, flush=True)
        k, lt, delta = _table(rng, N, L)
        tab = gpu.PerturbTable(k, lt, delta, ["d_cdm", "d_b", "t_tot"])
        ot = oracle.Transfer(k, lt, delta)
        f = _fbox(rng, N)
        r, w = gpu.Grid(N, L), gpu.Grid(N, L)
        r.set_fbox(f)
        r.momentum_space = 1
        for fi, ltv in ((0, 0.3), (2, lt[0] - 1.0), (1, lt[-1])):
            sp = tab.params(fi, ltv)
            ti, ut = ot.find_tau(ltv)
            assert (sp.tau_index, sp.u_tau) == (ti, ut)
            for inv in (False, True):
                name = "kernel_inv_transfer_function" if inv else "kernel_transfer_function"
                assert gpu.fft_apply_kernel_dg(w, r, name, sp) == 0
                ref = ot.apply(f, L, fi, ti, ut, invert=inv)
                err = rel(w.fbox(), ref)
                print(f"  f={fi} tau={ltv:.2f} inv={inv}: rel err {err:.1e}", flush=True)
                assert err < 1e-15
        # the host callbacks and the spline search agree with the oracle's
        for kk in (k[0], k[5], 0.5 * (k[10] + k[11]), k[-1], 2 * k[-1]):
            i1, u1 = ctypes.c_int(), ctypes.c_double()
            gpu.lib().perturbSplineFindK(ctypes.byref(tab.spline), kk, ctypes.byref(i1),
                                         ctypes.byref(u1))
            assert (i1.value, u1.value) == ot.find_k(kk)
            v = gpu.lib().perturbSplineInterp0(ctypes.byref(tab.spline), kk, 0.3, 1)
            ti, ut = ot.find_tau(0.3)
            assert v == ot.interp(i1.value, ti, u1.value, ut, 1)
        tab.close()


def test_generate_perturbation_grid_vs_oracle(gpu, oracle):
    """src/generate_grids.c:28-67: transfer kernel at log_tau_source, growth
    rescaling, c2r; an unknown title returns 1"""
    from mitos_amd import _lib
    rng = np.random.default_rng(34)
    N, L, seed = 32, 256.0, 5
    k, lt, delta = _table(rng, N, L)
    tab = gpu.PerturbTable(k, lt, delta, ["d_cdm", "d_b", "t_cdm"])
    ot = oracle.Transfer(k, lt, delta)
    cosmo = _lib.Cosmology()
    cosmo.log_tau_source = 0.45
    grf, grid = gpu.Grid(N, L), gpu.Grid(N, L)
    assert gpu.generate_complex_grf(grf, gpu.rand_uint64_init(seed)) == 0
    assert gpu.enforce_hermiticity(grf) == 0
    assert gpu.generatePerturbationGrid(cosmo, tab, grf, grid, "t_cdm", None, 0.8) == 0
    assert grid.momentum_space == 0
    f = oracle.grf(N, L, seed)
    ti, ut = ot.find_tau(0.45)
    g = ot.apply(f, L, 2, ti, ut)
    g = oracle.apply_kernel(g, L, oracle.K_CONSTANT, (0.8,))
    ref = oracle.fft_c2r(g) / L ** 3
    err = rel(grid.box()[:, :, :N], ref)
    print("case generatePerturbationGrid rel err", err, flush=True)
    assert err < 1e-13
    assert gpu.generatePerturbationGrid(cosmo, tab, grf, grid, "nope", None, 1.0) == 1
    tab.close()


def test_serial_array_helpers(gpu, oracle):
    """fft_apply_kernel(_float)(_float_mpi), fft_normalize_c2r,
    fft_normalize_r2c(_float)(_float_mpi), mitos_fft_c2r, the window
    corrections and solvePoisson on unpadded device arrays"""
    import torch
    lib = gpu.lib()
    rng = np.random.default_rng(35)
    N, L = 32, 120.0
    f = _fbox(rng, N)
    fd = torch.from_numpy(f).cuda()
    out = torch.empty_like(fd)
    from mitos_amd._lib import kernel_ptr
    torch.cuda.synchronize()
    assert lib.fft_apply_kernel(out.data_ptr(), fd.data_ptr(), N, L, kernel_ptr("kernel_dy"),
                                None) == 0
    assert rel(out.cpu().numpy(), oracle.apply_kernel(f, L, oracle.K_DY, ())) < 1e-15
    ff = fd.to(torch.complex64)
    of = torch.empty_like(ff)
    ref32 = oracle.apply_kernel(ff.cpu().numpy().astype(np.complex128), L, oracle.K_INV_POISSON,
                                ())
    torch.cuda.synchronize()
    assert lib.fft_apply_kernel_float(of.data_ptr(), ff.data_ptr(), N, L,
                                      kernel_ptr("kernel_inv_poisson"), None) == 0
    # the product in double, rounded once to float (as C does)
    assert np.array_equal(of.cpu().numpy(), ref32.astype(np.complex64))
    # the MPI variant over x-slabs [X0, X0 + NX) of the full grid
    X0, NX = 8, 12
    sl = ff[X0:X0 + NX].contiguous()
    osl = torch.empty_like(sl)
    torch.cuda.synchronize()
    assert lib.fft_apply_kernel_float_mpi(osl.data_ptr(), sl.data_ptr(), N, X0, NX, L,
                                          kernel_ptr("kernel_inv_poisson"), None) == 0
    assert np.array_equal(osl.cpu().numpy(), ref32.astype(np.complex64)[X0:X0 + NX])
    # normalisations
    g = fd.clone()
    torch.cuda.synchronize()
    assert lib.fft_normalize_r2c(g.data_ptr(), N, L) == 0
    vol = L * L * L
    assert np.array_equal(g.cpu().numpy(), f * (vol / N ** 3))
    g32 = ff.clone()
    torch.cuda.synchronize()
    assert lib.fft_normalize_r2c_float(g32.data_ptr(), N, L) == 0
    want32 = (ff.cpu().numpy().astype(np.complex128) * (vol / N ** 3)).astype(np.complex64)
    assert np.array_equal(g32.cpu().numpy(), want32)
    g32 = ff.clone()
    torch.cuda.synchronize()
    assert lib.fft_normalize_r2c_float_mpi(g32.data_ptr(), N, N * N * (N // 2 + 1), L) == 0
    assert np.array_equal(g32.cpu().numpy(), want32)
    real = torch.from_numpy(rng.standard_normal((N, N, N))).cuda()
    r2 = real.clone()
    torch.cuda.synchronize()
    assert lib.fft_normalize_c2r(r2.data_ptr(), N, L) == 0
    assert np.array_equal(r2.cpu().numpy(), real.cpu().numpy() / vol)
    # mitos_fft_c2r (unnormalised, numpy c2r semantics)
    back = torch.empty((N, N, N), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert lib.mitos_fft_c2r(N, fd.data_ptr(), back.data_ptr()) == 0
    assert rel(back.cpu().numpy(), oracle.fft_c2r(f)) < 1e-14
    # window corrections
    for name, order, fp32 in (("undoCICWindow", 2, False), ("undoTSCWindow", 3, False),
                              ("undoCICWindowFloat", 2, True), ("undoTSCWindowFloat", 3, True)):
        src = (ff if fp32 else fd).clone()
        torch.cuda.synchronize()
        assert getattr(lib, name)(src.data_ptr(), N, L) == 0
        base = (ff if fp32 else fd).cpu().numpy().astype(np.complex128)
        ref = oracle.apply_kernel(base, L, oracle.K_UNDO_WINDOW, (order, N, L))
        if fp32:
            assert np.array_equal(src.cpu().numpy(), ref.astype(np.complex64)), name
        else:
            assert rel(src.cpu().numpy(), ref) < 1e-15, name
    # solvePoisson (src/poisson.c:29-55): phi with Laplacian f, DC kept
    rho = rng.standard_normal((N, N, N))
    fr_ = torch.from_numpy(rho).cuda()
    phi = torch.empty_like(fr_)
    torch.cuda.synchronize()
    assert lib.solvePoisson(phi.data_ptr(), fr_.data_ptr(), N, L) == 0
    F = oracle.fft_r2c(rho) * (L ** 3 / N ** 3)
    F = oracle.apply_kernel(F, L, oracle.K_INV_POISSON, ())
    ref = oracle.fft_c2r(F) / L ** 3
    err = rel(phi.cpu().numpy(), ref)
    print("case solvePoisson rel err", err, flush=True)
    assert err < 1e-13


def test_powerspec_float_mpi(gpu, oracle):
    """calc_cross_powerspec_float_mpi (src/calc_powerspec.c:152-217) on one
    rank: the whole grid as one slab, long counts"""
    import torch
    from mitos_amd import _lib
    rng = np.random.default_rng(36)
    N, L, bins = 32, 256.0, 16
    a = _fbox(rng, N).astype(np.complex64)
    b = _fbox(rng, N).astype(np.complex64)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    k = np.empty(bins)
    P = np.empty(bins)
    obs = np.zeros(bins, np.int64)
    torch.cuda.synchronize()
    gpu.lib().calc_cross_powerspec_float_mpi(N, 0, N, L, da.data_ptr(), db.data_ptr(), bins,
                                             k.ctypes.data_as(_lib.c_double_p),
                                             P.ctypes.data_as(_lib.c_double_p),
                                             obs.ctypes.data_as(_lib.c_long_p))
    assert gpu.lib().mitos_last_error() == 0
    rk, rP, robs = oracle.powerspec(a.astype(np.complex128), b.astype(np.complex128), L, bins)
    assert np.array_equal(obs, robs.astype(np.int64))
    ok = robs > 0
    assert np.allclose(P[ok], rP[ok], rtol=1e-12, atol=0)
    assert np.allclose(k[ok], rk[ok], rtol=1e-13, atol=0)


def test_sample_norm_and_grid_tsc(gpu, oracle):
    from mitos_amd import _lib
    lib = gpu.lib()
    st = lib.rand_uint64_init(42)
    got = np.array([lib.sampleNorm(ctypes.byref(st)) for _ in range(512)])
    ref = oracle.normals(42, 0, 512)
    assert np.array_equal(got, ref)  # the same host libm as the oracle
    st = lib.rand_uint64_init(7)
    u = lib.sampleUniform(ctypes.byref(st))
    assert u == (float(oracle.draws(7, 0, 1)[0]) + 0.5) / 2.0 ** 64
    # gridTSC_dg at single points of a device grid
    N, L = 16, 40.0
    rng = np.random.default_rng(37)
    box = np.zeros((N, N, N + 2))
    box[:, :, :N] = rng.standard_normal((N, N, N))
    g = gpu.Grid(N, L)
    g.set_box(box)
    lrs = _lib.LeftRightSlice(None, g.s.box, None, 0, 0, N, 0, 0, 0)
    for x, y, z in ((0.0, 0.0, 0.0), (39.99, 0.01, 17.3), (-3.2, 45.0, 2.5), (10.0, 20.0, 30.0)):
        v = lib.gridTSC_dg(ctypes.byref(lrs), x, y, z, L, N)
        assert v == oracle.grid_tsc(box, L, x, y, z), (x, y, z)


def test_displace_velocity_vs_oracle(gpu, oracle):
    """src/mitos.c:824-908 then :1072-1077 in one batched call"""
    import torch
    N, L = 32, 256.0
    rng = np.random.default_rng(38)
    lib = gpu.lib()
    psi = np.zeros((3, N, N, N + 2))
    psi[:, :, :, :N] = 3.0 * rng.standard_normal((3, N, N, N))
    vel = np.zeros((3, N, N, N + 2))
    vel[:, :, :, :N] = 100.0 * rng.standard_normal((3, N, N, N))
    grids = []
    for arr in list(psi) + list(vel):
        g = gpu.Grid(N, L)
        g.set_box(arr)
        g.momentum_space = 0
        grids.append(g)
    n = N ** 3
    X, Y, Z, _ = oracle.lattice(N, L)
    dP = [torch.from_numpy(a.copy()).cuda() for a in (X, Y, Z)]
    dV = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    assert lib.mitos_displace_velocity_tsc_dg(*[g.ref() for g in grids], n,
                                              *[t.data_ptr() for t in dP],
                                              *[t.data_ptr() for t in dV]) == 0
    V = oracle.displace_velocity_tsc(psi, vel, L, X, Y, Z)
    for t, r in zip(dP + dV, [X, Y, Z] + V):
        got = t.cpu().numpy()
        same = float(np.mean(got == r))
        print("case displace+velocity: identical fraction", same, flush=True)
        assert same > 0.999
        ulp = np.spacing(np.abs(r).astype(np.float32))
        assert np.all(np.abs(got.astype(np.float64) - r) <= ulp)
    # velocities only, at the lattice (no moves)
    X, Y, Z, _ = oracle.lattice(N, L)
    dP = [torch.from_numpy(a.copy()).cuda() for a in (X, Y, Z)]
    torch.cuda.synchronize()
    assert lib.mitos_displace_velocity_tsc_dg(None, None, None, *[g.ref() for g in grids[3:]],
                                              n, *[t.data_ptr() for t in dP],
                                              *[t.data_ptr() for t in dV]) == 0
    V = oracle.displace_velocity_tsc(None, vel, L, X, Y, Z)
    for t, r in zip(dV, V):
        assert np.array_equal(t.cpu().numpy(), r)


def test_particles_export_layout(gpu):
    """src/mitos.c:1097-1107: interleaved double coordinates and velocities,
    masses, ids = index + first_id"""
    import torch
    rng = np.random.default_rng(39)
    n = (1 << 20) + 12345  # more than one staging pass
    P = [rng.uniform(0, 100, n).astype(np.float32) for _ in range(3)]
    V = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    dP = [torch.from_numpy(a).cuda() for a in P]
    dV = [torch.from_numpy(a).cuda() for a in V]
    coords = np.empty((n, 3))
    vels = np.empty((n, 3))
    masses = np.empty(n)
    ids = np.empty(n, np.int64)
    torch.cuda.synchronize()
    assert gpu.lib().mitos_particles_export(n, *[t.data_ptr() for t in dP],
                                            *[t.data_ptr() for t in dV], 2.5, 7,
                                            coords.ctypes.data, vels.ctypes.data,
                                            masses.ctypes.data, ids.ctypes.data) == 0
    assert np.array_equal(coords, np.stack(P, 1).astype(np.float64))
    assert np.array_equal(vels, np.stack(V, 1).astype(np.float64))
    assert np.all(masses == 2.5)
    assert np.array_equal(ids, np.arange(n) + 7)


def test_shrink_grid_vs_oracle(gpu, oracle):
    import torch
    rng = np.random.default_rng(40)
    N, M, L = 48, 12, 100.0
    box = np.zeros((N, N, N + 2))
    box[:, :, :N] = rng.standard_normal((N, N, N))
    g = gpu.Grid(N, L)
    g.set_box(box)
    out = torch.empty((M, M, M), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert gpu.lib().shrinkGrid_dg(out.data_ptr(), g.ref(), M, N) == 0
    ref = oracle.shrink_grid(box[:, :, :N], M)
    assert np.array_equal(out.cpu().numpy(), ref)
    flat = torch.from_numpy(box[:, :, :N].copy()).cuda()
    out2 = torch.empty_like(out)
    torch.cuda.synchronize()
    assert gpu.lib().shrinkGrid(out2.data_ptr(), flat.data_ptr(), M, N) == 0
    assert np.array_equal(out2.cpu().numpy(), ref)
    assert gpu.lib().shrinkGrid(out2.data_ptr(), flat.data_ptr(), 7, N) == 1


def test_field_file_round_trip(gpu):
    """writeFieldFile_dg / readFieldFile_dg (padded N x N x (N+2) dataset),
    writeFieldFile / readFieldFile (N^3), and a padded file read back to the
    host without its padding"""
    from mitos_amd import _lib
    lib = gpu.lib()
    N, L = 24, 75.0
    rng = np.random.default_rng(41)
    box = np.zeros((N, N, N + 2))
    box[:, :, :N] = rng.standard_normal((N, N, N))
    g = gpu.Grid(N, L)
    g.set_box(box)
    g.momentum_space = 0
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "field.hdf5").encode()
        rc = lib.writeFieldFile_dg(g.ref(), fn)
        if rc == 4:
            pytest.skip("libhdf5 is not available on this machine")
        assert rc == 0
        h = gpu.Grid(N, L)
        h.momentum_space = 1
        assert lib.readFieldFile_dg(h.ref(), fn) == 0
        assert h.momentum_space == 0
        assert np.array_equal(h.box(), box)
        # the host reader drops the padding
        pb = _lib.c_double_p()
        n = ctypes.c_int()
        bl = ctypes.c_double()
        assert lib.readFieldFile(ctypes.byref(pb), ctypes.byref(n), ctypes.byref(bl), fn) == 0
        assert n.value == N and bl.value == L
        host = np.ctypeslib.as_array(pb, shape=(N, N, N)).copy()
        ctypes.CDLL(None).free(pb)
        assert np.array_equal(host, box[:, :, :N])
        # an unpadded file into a grid
        fn2 = os.path.join(d, "plain.hdf5").encode()
        src = np.ascontiguousarray(box[:, :, :N])
        assert lib.writeFieldFile(src.ctypes.data, N, L, fn2) == 0
        h2 = gpu.Grid(N, L)
        assert lib.readFieldFile_dg(h2.ref(), fn2) == 0
        assert np.array_equal(h2.box()[:, :, :N], src)
        # momentum space: not exported (src/output_mpi.c:79-82)
        g.momentum_space = 1
        assert lib.writeFieldFile_dg(g.ref(), fn) == 1


def test_monge_ampere_vs_oracle(gpu, oracle):
    """solveMongeAmpere (src/monge_ampere.c:50-183): three cycles on a smooth
    density of amplitude 0.2 against the oracle, the per-cycle residuals, and
    the momentum-space check"""
    N, L = 32, 128.0
    rng = np.random.default_rng(42)
    f = oracle.apply_kernel(_fbox(rng, N), L, oracle.K_GAUSSIAN, (6.0,))
    delta = oracle.fft_c2r(f)
    delta *= 0.2 / np.abs(delta).max()
    F = oracle.fft_r2c(delta) * (L ** 3 / N ** 3)
    od, op, ow = oracle.Grid(N, L), oracle.Grid(N, L), oracle.Grid(N, L)
    od.box[:, :, :N] = delta
    od.fbox[:] = F
    od.s.momentum_space = b"\x01"
    rc, eps = oracle.solve_monge_ampere(op, od, ow, 3)
    assert rc == 0
    d, p, w = gpu.Grid(N, L), gpu.Grid(N, L), gpu.Grid(N, L)
    box = np.zeros((N, N, N + 2))
    box[:, :, :N] = delta
    d.set_box(box)
    d.set_fbox(F)
    d.momentum_space = 1
    assert gpu.lib().solveMongeAmpere(p.ref(), d.ref(), w.ref(), 3) == 0
    assert p.momentum_space == 1
    err = rel(p.fbox(), op.fbox)
    got = np.zeros(3)
    n = gpu.lib().mitos_monge_ampere_residuals(got.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                3)
    print("case Monge-Ampere: potential rel err", err, "residuals", got, eps, flush=True)
    assert n == 3
    assert err < 1e-12
    assert np.all(np.abs(got / eps - 1) < 1e-9)
    assert np.all(np.diff(got) < 0)
    d.momentum_space = 0
    assert gpu.lib().solveMongeAmpere(p.ref(), d.ref(), w.ref(), 1) == 1


def test_ngeniclike_grf_vs_oracle(gpu, oracle):
    """generate_ngeniclike_grf on the device against the oracle's loop-by-loop
    restatement: the same RANLUX streams (exact), host-libm against device
    sin/cos/log (last bits); a slab [X0, X0+NX) is the matching part of the
    whole grid, conjugate partners from other slabs included"""
    import torch
    lib = gpu.lib()
    for N, L, seed in ((16, 100.0, 42), (48, 333.0, 7), (64, 256.0, 123456)):
        print("case ngeniclike", N, L, seed, flush=True)
        ref = oracle.ngeniclike_grf(N, L, seed)
        out = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        assert lib.generate_ngeniclike_grf(out.data_ptr(), N, N, 0, 0, L, seed) == 0
        got = out.cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        zeros_same = np.array_equal(got == 0, ref == 0)
        print(f"  max err / max {err:.1e}, identical fraction {np.mean(got == ref):.4f}",
              flush=True)
        assert zeros_same
        assert err < 1e-15
        X0, NX = N // 4, N // 2
        sl = torch.empty((NX, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        assert lib.generate_ngeniclike_grf(sl.data_ptr(), N, NX, X0, 0, L, seed) == 0
        assert np.array_equal(sl.cpu().numpy(), got[X0:X0 + NX])
