"""GPU parity of the whole IC pipeline (fused FP64/FP32 and the reference-order
FP64 plain path) against the oracle's reference-order FP64 run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, N, L, seed, bins, run2lpt=True):
    return oracle.ic_run(N, L, seed, bins=bins, run2lpt=run2lpt, want_psi=True,
                         want_density=True)


def _pk_rel(P, ref, obs):
    ok = obs > 1
    return np.abs(P[ok] - ref[ok]) / np.abs(ref[ok])


def _pos_err(X, Xr, L):
    d = np.abs(X.astype(np.float64) - Xr.astype(np.float64))
    d = np.minimum(d, L - d)  # periodic
    return d.max()


@pytest.mark.parametrize("N,L,seed,bins,run2lpt", [
    (32, 128.0, 42, 16, True), (64, 256.0, 42, 32, True), (32, 256.0, 3, 16, False),
    (24, 96.0, 5, 12, True), (48, 150.0, 6, 16, True), (96, 300.0, 7, 24, True)])
def test_fused_fp64_vs_oracle(gpu, oracle, N, L, seed, bins, run2lpt):
    print("case fused fp64", N, L, seed, bins, run2lpt, flush=True)
    ref = _oracle_run(oracle, N, L, seed, bins, run2lpt)
    cfg = gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, run2lpt=int(run2lpt),
                       precision=gpu.MITOS_FP64)
    ic = gpu.IC(cfg)
    out = ic.run()
    assert np.array_equal(out["obs"], ref["obs"])
    err = _pk_rel(out["P"], ref["P"], ref["obs"])
    print("  P(k) max rel err", err.max(), flush=True)
    assert err.max() < 1e-10
    psi = ic.displacements().cpu().numpy()
    rpsi = ref["psi"][:, :, :, :N]
    e = np.linalg.norm(psi - rpsi) / np.linalg.norm(rpsi)
    print("  psi rel err", e, flush=True)
    assert e < 1e-12
    X, Y, Z = (t.cpu().numpy() for t in ic.positions())
    pe = max(_pos_err(X, ref["X"], L), _pos_err(Y, ref["Y"], L), _pos_err(Z, ref["Z"], L))
    print("  max position diff", pe, flush=True)
    # the reference's 27-tap sums in its order: the same float positions
    assert np.array_equal(X, ref["X"]) and np.array_equal(Y, ref["Y"])
    assert np.array_equal(Z, ref["Z"])
    dens = ic.density().cpu().numpy()
    de = np.linalg.norm(dens - ref["density"]) / np.linalg.norm(ref["density"])
    print("  density rel err", de, flush=True)
    assert de < 1e-11  # measured <= 8.2e-13
    ids = ic.ids().cpu().numpy()
    assert np.array_equal(ids, ref["id"])
    ic.close()


@pytest.mark.parametrize("N,L,seed,bins", [
    (32, 128.0, 42, 16), (64, 256.0, 42, 32), (48, 192.0, 8, 16), (96, 300.0, 9, 24),
    (32, 100.0 / 3.0, 10, 16), (48, 100.0, 11, 16)])
def test_fused_fp32_vs_oracle(gpu, oracle, N, L, seed, bins):
    print("case fused fp32", N, L, seed, bins, flush=True)
    ref = _oracle_run(oracle, N, L, seed, bins)
    cfg = gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, precision=gpu.MITOS_FP32)
    ic = gpu.IC(cfg)
    out = ic.run()
    assert np.array_equal(out["obs"], ref["obs"])
    err = _pk_rel(out["P"], ref["P"], ref["obs"])
    print("  P(k) max rel err (fp32)", err.max(), flush=True)
    assert err.max() < 5e-6  # measured <= 5.1e-7
    psi = ic.displacements().cpu().numpy().astype(np.float64)
    rpsi = ref["psi"][:, :, :, :N]
    e = np.linalg.norm(psi - rpsi) / np.linalg.norm(rpsi)
    print("  psi rel err (fp32)", e, flush=True)
    assert e < 4e-6  # measured <= 4e-7
    # positions: lattice-stencil moves where the lattice sits on grid nodes,
    # 27-tap moves at the actual float positions otherwise (L = 100/3, 100/48)
    X, Y, Z = (t.cpu().numpy() for t in ic.positions())
    pe = max(_pos_err(X, ref["X"], L), _pos_err(Y, ref["Y"], L), _pos_err(Z, ref["Z"], L))
    print("  max position diff / cell (fp32)", pe * N / L, flush=True)
    assert pe < 1e-4 * L / N
    ic.close()


def test_reference_order_fp64_vs_oracle(gpu, oracle):
    N, L, seed, bins = 32, 256.0, 42, 16
    print("case reference order", N, flush=True)
    ref = _oracle_run(oracle, N, L, seed, bins)
    out = gpu.run_reference_order(gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins),
                                  positions=True)
    assert np.array_equal(out["obs"], ref["obs"])
    err = _pk_rel(out["P"], ref["P"], ref["obs"])
    print("  P(k) max rel err", err.max(), flush=True)
    assert err.max() < 1e-10
    for c in "XYZ":
        got = out[c].cpu().numpy()
        ulp = np.spacing(np.abs(ref[c]).astype(np.float32)).astype(np.float64)
        d = np.abs(got.astype(np.float64) - ref[c].astype(np.float64))
        print(f"  {c}: identical fraction {np.mean(got == ref[c])}", flush=True)
        assert np.all(d <= ulp)


@pytest.mark.parametrize("N,M,L,prec", [
    (32, 24, 128.0, "fp64"), (32, 48, 100.0, "fp64"), (64, 40, 256.0, "fp64"),
    (32, 48, 100.0, "fp32"), (64, 40, 256.0, "fp32")])
def test_particle_lattice_other_than_grid(gpu, oracle, N, M, L, prec):
    # [ParticleType_0] CubeRootNumber M != GridSize N (src/particle.c:87-129):
    # M^3 particles on the lattice of spacing (float)L / M, moved by the
    # 27-tap TSC at their positions and deposited with CIC, the overdensity
    # normalised by the M^3 particles (analyse/src/mitos_read.c:363-376)
    print("case lattice", N, M, L, prec, flush=True)
    seed, bins = 11, 16
    ref = oracle.ic_run(N, L, seed, bins=bins, want_density=True, M=M)
    p = gpu.MITOS_FP64 if prec == "fp64" else gpu.MITOS_FP32
    cfg = gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, precision=p, M=M)
    ic = gpu.IC(cfg)
    out = ic.run()
    assert np.array_equal(out["obs"], ref["obs"])
    err = _pk_rel(out["P"], ref["P"], ref["obs"])
    X, Y, Z = (t.cpu().numpy() for t in ic.positions())
    assert X.size == M ** 3
    pe = max(_pos_err(X, ref["X"], L), _pos_err(Y, ref["Y"], L), _pos_err(Z, ref["Z"], L))
    dens = ic.density().cpu().numpy()
    de = np.linalg.norm(dens - ref["density"]) / np.linalg.norm(ref["density"])
    print(f"  P(k) max rel err {err.max():.2e}, max position diff {pe:.2e}, "
          f"density rel err {de:.2e}", flush=True)
    if prec == "fp64":
        assert err.max() < 1e-10
        assert np.array_equal(X, ref["X"]) and np.array_equal(Y, ref["Y"])
        assert np.array_equal(Z, ref["Z"])
        assert de < 1e-11
    else:
        assert err.max() < 5e-6
        assert pe < 1e-4 * L / N
        assert de < 1e-5
    assert np.array_equal(ic.ids().cpu().numpy(), ref["id"])
    if M < N:
        # more particles than the context was made for: MITOS_ERR_DIMS
        from mitos_amd._lib import MitosError
        with pytest.raises(MitosError) as e:
            ic.run(gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, precision=p, M=N + 8))
        assert e.value.code == 1
    ic.close()
    if prec == "fp64" and M < N:
        # ensembles keep the lattice on the grid
        from mitos_amd._lib import MitosError
        with pytest.raises(MitosError) as e:
            gpu.Ensemble([cfg])
        assert e.value.code == 4
    if prec == "fp64":
        # the reference call order through the drop-in functions
        rp = gpu.run_reference_order(cfg, positions=True)
        assert np.array_equal(rp["obs"], ref["obs"])
        assert _pk_rel(rp["P"], ref["P"], ref["obs"]).max() < 1e-10
        assert rp["X"].numel() == M ** 3


def test_fused_repeatable(gpu):
    cfg = gpu.ICConfig(N=32, boxlen=100.0, seed=5, bins=16, precision=gpu.MITOS_FP32)
    ic = gpu.IC(cfg)
    a = ic.run()
    b = ic.run()
    ok = a["obs"] > 0
    assert np.array_equal(a["P"][ok], b["P"][ok])
    ic.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_every_fft_schedule_variant(gpu, oracle, variant):
    # each schedule the autotuner can pick gives the oracle's P(k) and psi
    N, L, seed, bins = 32, 128.0, 11, 16
    print("case schedule variant", variant, flush=True)
    ref = _oracle_run(oracle, N, L, seed, bins)
    gpu.set_fft_schedule(N, gpu.MITOS_FP64, [variant] * 4)
    ic = gpu.IC(gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, precision=gpu.MITOS_FP64))
    assert all(v == variant for v, _ in ic.schedules().values())
    out = ic.run()
    err = _pk_rel(out["P"], ref["P"], ref["obs"])
    assert err.max() < 1e-10
    dens = ic.density().cpu().numpy()
    de = np.linalg.norm(dens - ref["density"]) / np.linalg.norm(ref["density"])
    assert de < 1e-11
    ic.close()
    gpu.set_fft_schedule(N, gpu.MITOS_FP32, [variant] * 4)
    ic = gpu.IC(gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, precision=gpu.MITOS_FP32))
    out = ic.run()
    assert _pk_rel(out["P"], ref["P"], ref["obs"]).max() < 5e-6
    ic.close()


@pytest.mark.parametrize("N,L,sigma8", [(128, 250.0, 0.81), (256, 500.0, 0.81), (128, 64.0, 3.0)])
def test_two_pass_deposit_matches_atomic(gpu, N, L, sigma8, monkeypatch):
    # the FP32 lattice deposit through window slots + gather (N >= 128) and the
    # one-pass deposit with atomic window flushes sum the same integers: the
    # overdensity is bit-identical; the last case (small cells, large
    # amplitude) sends many particles outside their tile's window
    print("case two-pass deposit", N, L, sigma8, flush=True)
    cfg = gpu.ICConfig(N=N, boxlen=L, seed=9, bins=32, sigma8=sigma8, precision=gpu.MITOS_FP32)
    ic = gpu.IC(cfg)
    a = ic.run()
    da = ic.density().cpu().numpy()
    ic.close()
    monkeypatch.setenv("MITOS_CIC_ATOMIC", "1")
    ic = gpu.IC(cfg)
    b = ic.run()
    db = ic.density().cpu().numpy()
    ic.close()
    assert np.array_equal(da, db)
    ok = a["obs"] > 0
    assert np.array_equal(a["P"][ok], b["P"][ok])
    assert abs(float(da.astype(np.float64).mean())) < 1e-6


@pytest.mark.parametrize("N,L,sigma8", [(128, 250.0, 0.81), (256, 500.0, 0.81),
                                         (128, 64.0, 3.0)])
def test_staged_displacement_matches_direct(gpu, N, L, sigma8, monkeypatch):
    # the FP32 displacement over tiles with its taps staged in LDS, the same
    # kernel with every tap read from the grids, and the one-wave row kernel
    # give the same positions, bit for bit (the last case spreads the taps
    # beyond the staging windows, so the staged and direct paths both run)
    print("case staged displacement", N, L, sigma8, flush=True)
    cfg = gpu.ICConfig(N=N, boxlen=L, seed=4, bins=16, sigma8=sigma8, precision=gpu.MITOS_FP32)
    pos = []
    for env in ({}, {"MITOS_TSC_DIRECT": "1"}, {"MITOS_TSC_KERNEL": "wave"}):
        for k in ("MITOS_TSC_DIRECT", "MITOS_TSC_KERNEL"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ic = gpu.IC(cfg)
        ic.run()
        pos.append([t.cpu().numpy() for t in ic.positions()])
        ic.close()
    for other in pos[1:]:
        for a, b in zip(pos[0], other):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("N,L,sigma8", [(64, 100.0, 0.81), (128, 250.0, 0.81),
                                         (256, 500.0, 0.81), (128, 64.0, 3.0)])
def test_fp64_tiled_displacement_matches_per_particle(gpu, N, L, sigma8, monkeypatch):
    # the FP64 moves in the reference's 27-tap arithmetic with the taps staged
    # in LDS per 8x8x16 tile, and the per-particle kernel reading every tap
    # from the grids, give the same positions bit for bit (the last case
    # spreads the taps beyond the windows, so both of the tile kernel's
    # paths run)
    print("case fp64 tiled displacement", N, L, sigma8, flush=True)
    cfg = gpu.ICConfig(N=N, boxlen=L, seed=6, bins=16, sigma8=sigma8, precision=gpu.MITOS_FP64)
    pos = []
    for env in ({}, {"MITOS_TSC_KERNEL": "wave"}):
        monkeypatch.delenv("MITOS_TSC_KERNEL", raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ic = gpu.IC(cfg)
        ic.run()
        pos.append([t.cpu().numpy() for t in ic.positions()])
        ic.close()
    for a, b in zip(pos[0], pos[1]):
        assert np.array_equal(a, b)


def test_fp32_pk_vs_fp64_256(gpu):
    # the FP32 pipeline at the cell size of the 1024^3 / 2000 Mpc/h workload:
    # P(k) within 1e-5 of the FP64 pipeline in every bin with more than one mode
    print("case fp32 vs fp64 P(k) 256", flush=True)
    out = {}
    for prec in (gpu.MITOS_FP64, gpu.MITOS_FP32):
        cfg = gpu.ICConfig(N=256, boxlen=500.0, seed=42, bins=64, precision=prec)
        ic = gpu.IC(cfg)
        out[prec] = ic.run()
        ic.close()
    a, b = out[gpu.MITOS_FP32], out[gpu.MITOS_FP64]
    err = _pk_rel(a["P"], b["P"], b["obs"])
    print("  max", err.max(), "rms", np.sqrt((err ** 2).mean()), flush=True)
    assert err.max() < 1e-5


def test_every_fft_schedule_variant_128(gpu, oracle):
    # at 128^3 the variants differ in tile widths and rows per work-group as
    # at the large sizes; each gives the oracle's P(k)
    N, L, seed, bins = 128, 250.0, 13, 24
    ref = _oracle_run(oracle, N, L, seed, bins)
    ok = ref["obs"] > 1
    for variant in range(4):
        print("case schedule variant 128", variant, flush=True)
        for prec in (gpu.MITOS_FP64, gpu.MITOS_FP32):
            gpu.set_fft_schedule(N, prec, [variant] * 4)
            ic = gpu.IC(gpu.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, precision=prec))
            P = ic.run()["P"]
            ic.close()
            err = _pk_rel(P, ref["P"], ref["obs"])
            nrel = np.linalg.norm(P[ok] - ref["P"][ok]) / np.linalg.norm(ref["P"][ok])
            if prec == gpu.MITOS_FP64:
                # float positions: a last-bit difference of psi moves a few
                # particles by one float ulp (the FP64 contract is norm-relative)
                assert nrel < 1e-10 and err.max() < 1e-6
            else:
                assert err.max() < 5e-6


def test_enqueue_on_caller_stream_then_run(gpu):
    """A realization enqueued on a caller stream, then another run on the
    context's own stream with a different seed, then the first again on the
    caller stream: each waits for the previous one on the device, so every
    P(k) equals that of a separate run."""
    import torch
    N, L, bins = 64, 256.0, 16
    cfg_a = gpu.ICConfig(N=N, boxlen=L, seed=21, bins=bins, precision=gpu.MITOS_FP32)
    cfg_b = gpu.ICConfig(N=N, boxlen=L, seed=22, bins=bins, precision=gpu.MITOS_FP32)
    ic = gpu.IC(cfg_a)
    ref_a = ic.run(cfg_a)
    ref_b = ic.run(cfg_b)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    ic.enqueue(cfg_a, stream=s.cuda_stream)
    got_b = ic.run(cfg_b)
    ic.enqueue(cfg_a, stream=s.cuda_stream)
    got_a = ic.fetch_pk()
    ok = ref_a["obs"] > 0
    assert np.array_equal(got_b["P"][ok], ref_b["P"][ok])
    assert np.array_equal(got_a["P"][ok], ref_a["P"][ok])
    ic.close()


def test_two_contexts_on_two_streams(gpu):
    """Two pipelines in flight at once on two streams share only read-only
    tables: each realization equals a separate run bit for bit (P(k),
    positions, density)."""
    import torch
    N, L, bins = 256, 500.0, 32
    cfg_a = gpu.ICConfig(N=N, boxlen=L, seed=31, bins=bins, precision=gpu.MITOS_FP32)
    cfg_b = gpu.ICConfig(N=N, boxlen=L, seed=32, bins=bins, precision=gpu.MITOS_FP32)
    print("case two contexts", N, flush=True)
    refs = []
    for cfg in (cfg_a, cfg_b):
        ic = gpu.IC(cfg)
        out = ic.run(cfg)
        refs.append((out, [t.clone() for t in ic.positions()], ic.density().clone()))
        ic.close()
    c1, c2 = gpu.IC(cfg_a), gpu.IC(cfg_b)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for _ in range(3):
        c1.enqueue(cfg_a, stream=s1.cuda_stream)
        c2.enqueue(cfg_b, stream=s2.cuda_stream)
    for c, (ref, pos, dens) in zip((c1, c2), refs):
        got = c.fetch_pk()
        ok = ref["obs"] > 0
        assert np.array_equal(got["P"][ok], ref["P"][ok])
        assert all(torch.equal(a, b) for a, b in zip(c.positions(), pos))
        assert torch.equal(c.density(), dens)
        c.close()
