"""GPU: the hand-written adjoint (dL/d white noise through GRF, 2LPT, TSC
displacement, CIC and P(k)) against central finite differences of the FP64
forward, and the FP32 gradient against the FP64 one."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _modes(N, n, seed):
    """n noise modes: mostly inside the half spectrum, a few on the z' = 0
    plane (kept and overwritten by the Hermitian symmetry)"""
    rng = np.random.default_rng(seed)
    H = N // 2 + 1
    out = [(int(rng.integers(N)), int(rng.integers(N)), int(rng.integers(1, H - 1)))
           for _ in range(n - 8)]
    out += [(int(rng.integers(N // 2 + 1, N)), int(rng.integers(N)), 0) for _ in range(4)]
    out += [(int(rng.integers(1, N // 2)), int(rng.integers(N)), 0) for _ in range(4)]
    return out


def _fd(loss_of, noise, modes, eps):
    """central differences along Re and Im of each mode"""
    out = []
    for (x, y, z) in modes:
        g = []
        for d in (eps, 1j * eps):
            p = noise.clone()
            p[x, y, z] += d
            lp = loss_of(p)
            p[x, y, z] -= 2 * d
            lm = loss_of(p)
            g.append((lp - lm) / (2 * eps))
        out.append(complex(g[0], g[1]))
    return np.array(out)


def _pk_setup(ic, N, L, bins, precision, seed=7, target_seed=8):
    from mitos_amd import adjoint
    cfg = ic.ICConfig(N=N, boxlen=L, seed=seed, bins=bins, precision=precision)
    ctx = ic.IC(cfg)
    tgt = ctx.noise_forward(cfg, ic.white_noise(N, target_seed, precision))["P"]
    noise = ic.white_noise(N, seed, precision)

    def loss_of(nz):
        return float(adjoint.pk_loss(torch.from_numpy(ctx.noise_forward(cfg, nz)["P"]), tgt))

    def grad_of(nz):
        n = nz.clone().requires_grad_(False)
        leaf = n.detach().requires_grad_(True)
        L_ = adjoint.pk_loss(adjoint.noise_power_spectrum(leaf, ctx, cfg), tgt)
        L_.backward()
        return leaf.grad, float(L_.detach())

    return cfg, ctx, noise, loss_of, grad_of


def test_noise_forward_matches_seed_run(gpu):
    ic = gpu
    N, L, bins = 32, 128.0, 16
    print("case noise forward vs seed run", flush=True)
    # the noise forward keeps positions in the working precision (the seed run
    # rounds them to float, as the reference does): agreement to that rounding
    for prec, tol in ((ic.MITOS_FP64, 1e-5), (ic.MITOS_FP32, 1e-4)):
        cfg = ic.ICConfig(N=N, boxlen=L, seed=42, bins=bins, precision=prec)
        ctx = ic.IC(cfg)
        a = ctx.run()
        b = ctx.noise_forward(cfg, ic.white_noise(N, 42, prec))
        ok = a["obs"] > 0
        assert np.array_equal(a["obs"], b["obs"])
        assert np.max(np.abs(a["P"][ok] - b["P"][ok]) / a["P"][ok]) < tol
        ctx.close()


@pytest.mark.parametrize("run2lpt", [0, 1])
def test_adjoint_vs_autograd_reference(gpu, run2lpt):
    # the FP64 gradient against torch.autograd through the same forward
    # restated on the CPU (oracle/adjoint_ref.py)
    from oracle import adjoint_ref as R
    from mitos_amd import adjoint
    ic = gpu
    N, L, bins = 16, 64.0, 8
    print("case adjoint vs autograd reference, run2lpt", run2lpt, flush=True)
    cfg = ic.ICConfig(N=N, boxlen=L, seed=7, bins=bins, run2lpt=run2lpt, precision=ic.MITOS_FP64)
    ctx = ic.IC(cfg)
    fw = R.Forward(N, L, bins, ic.sigma8_amplitude(0.965, 0.21, 0.81), run2lpt=bool(run2lpt))
    noise = ic.white_noise(N, 7, ic.MITOS_FP64)
    nz = noise.cpu().numpy()
    with torch.no_grad():
        P_ref = fw.power(fw.density(torch.tensor(nz))).numpy()
        tgt = fw.power(fw.density(torch.tensor(
            ic.white_noise(N, 8, ic.MITOS_FP64).cpu().numpy()))).numpy()
        dref = fw.density(torch.tensor(nz)).numpy()
    P = ctx.noise_forward(cfg, noise)["P"]
    ok = ~np.isnan(P_ref)
    print("  P(k) rel diff", np.max(np.abs(P[ok] - P_ref[ok]) / P_ref[ok]), flush=True)
    assert np.max(np.abs(P[ok] - P_ref[ok]) / P_ref[ok]) < 1e-10
    d = ctx.density().cpu().numpy()
    assert np.abs(d - dref).max() < 1e-9
    # P(k) loss
    _, g_ref = R.pk_gradient(fw, nz, tgt)
    leaf = noise.clone().requires_grad_(True)
    adjoint.pk_loss(adjoint.noise_power_spectrum(leaf, ctx, cfg), tgt).backward()
    g = leaf.grad.cpu().numpy()
    e = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print("  P(k) loss gradient rel err", e, flush=True)
    assert e < 1e-8
    # density loss
    r = np.random.default_rng(4).standard_normal((N, N, N))
    _, gd_ref = R.density_gradient(fw, nz, r)
    leaf = noise.clone().requires_grad_(True)
    dd = adjoint.noise_density(leaf, ctx, cfg)
    rt = torch.from_numpy(r).cuda()
    ((dd * rt).sum() + 0.5 * (dd * dd).sum()).backward()
    gd = leaf.grad.cpu().numpy()
    e = np.linalg.norm(gd - gd_ref) / np.linalg.norm(gd_ref)
    print("  density loss gradient rel err", e, flush=True)
    assert e < 1e-8
    ctx.close()


def test_noise_forward_conserves_mass(gpu):
    # every particle deposits weight 1: the overdensity has zero mean
    ic = gpu
    for prec in (ic.MITOS_FP32, ic.MITOS_FP64):
        N = 64
        cfg = ic.ICConfig(N=N, boxlen=200.0, seed=7, bins=16, precision=prec)
        ctx = ic.IC(cfg)
        ctx.noise_forward(cfg, ic.white_noise(N, 7, prec))
        d = ctx.density().double()
        print("case mass", prec, float(d.mean()), float(d.min()), flush=True)
        assert abs(float(d.mean())) < 1e-7
        assert float(d.min()) >= -1.0
        ctx.close()


def test_pk_adjoint_vs_finite_differences_fp64(gpu):
    ic = gpu
    N, L, bins = 32, 128.0, 16
    print("case P(k) adjoint FD fp64", flush=True)
    cfg, ctx, noise, loss_of, grad_of = _pk_setup(ic, N, L, bins, ic.MITOS_FP64)
    g, _ = grad_of(noise)
    g = g.cpu().numpy()
    modes = _modes(N, 64, 1)
    fd = _fd(loss_of, noise, modes, 1e-3)
    ga = np.array([g[m] for m in modes])
    scale = np.abs(ga).max()
    err = np.abs(fd - ga)
    print("  max |g|", scale, "max err", err.max(), flush=True)
    # particles crossing a CIC cell boundary within +-eps make the loss only
    # piecewise smooth: the differences agree to that level (the exact pin is
    # the autograd comparison above)
    assert np.all(err <= 2e-3 * scale + 2e-2 * np.abs(ga))
    # modes overwritten by the Hermitian planes have no influence
    for (x, y, z) in modes[-4:]:
        assert g[x, y, z] == 0
    ctx.close()


def test_density_adjoint_vs_finite_differences_fp64(gpu):
    ic = gpu
    from mitos_amd import adjoint
    N, L = 32, 100.0
    print("case density adjoint FD fp64", flush=True)
    cfg = ic.ICConfig(N=N, boxlen=L, seed=3, bins=8, precision=ic.MITOS_FP64)
    ctx = ic.IC(cfg)
    noise = ic.white_noise(N, 3, ic.MITOS_FP64)
    r = torch.from_numpy(np.random.default_rng(5).standard_normal((N, N, N))).cuda()

    def loss_of(nz):
        d = adjoint.noise_density(nz, ctx, cfg)
        return float((d * r).sum() + 0.5 * (d * d).sum())

    leaf = noise.clone().requires_grad_(True)
    d = adjoint.noise_density(leaf, ctx, cfg)
    ((d * r).sum() + 0.5 * (d * d).sum()).backward()
    g = leaf.grad.cpu().numpy()
    modes = _modes(N, 24, 2)
    fd = _fd(loss_of, noise, modes, 1e-3)
    ga = np.array([g[m] for m in modes])
    err = np.abs(fd - ga)
    print("  max |g|", np.abs(ga).max(), "max err", err.max(), flush=True)
    assert np.all(err <= 2e-2 * np.abs(ga).max() + 2e-2 * np.abs(ga))
    ctx.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_adjoint_is_bit_reproducible(gpu, prec):
    """the particle adjoint scatters in fixed point: two backward runs give
    the same gradient bits"""
    ic = gpu
    p = ic.MITOS_FP32 if prec == "fp32" else ic.MITOS_FP64
    print("case adjoint reproducible", prec, flush=True)
    _, ctx, noise, _, grad_of = _pk_setup(ic, 64, 200.0, 16, p)
    g1, l1 = grad_of(noise)
    g2, l2 = grad_of(noise)
    assert l1 == l2
    assert torch.equal(g1, g2)
    ctx.close()


def test_config5_fp32_adjoint_256(gpu):
    # 256^3, box 500 Mpc/h, seed 7, target P(k) of seed 8, 32 bins, w_b = 1/b;
    # the FP32 gradient against the FP64 one (whole field) and against FP64
    # central differences on 64 modes (eps = 1e-3)
    ic = gpu
    N, L, bins = 256, 500.0, 32
    print("case config5 fp32 adjoint", flush=True)
    _, ctx64, noise64, loss64, grad64 = _pk_setup(ic, N, L, bins, ic.MITOS_FP64)
    g64, l64 = grad64(noise64)
    _, ctx32, noise32, _, grad32 = _pk_setup(ic, N, L, bins, ic.MITOS_FP32)
    g32, l32 = grad32(noise32)
    print("  loss fp64", l64, "fp32", l32, flush=True)
    assert abs(l32 / l64 - 1) < 5e-5  # measured 5.3e-6
    a, b = g32.to(torch.complex128), g64
    rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    print("  fp32 vs fp64 gradient rel err", rel, flush=True)
    assert rel < 1e-4  # measured 9.7e-6
    modes = _modes(N, 64, 3)
    fd = _fd(loss64, noise64, modes, 1e-3)
    gb = b.cpu().numpy()
    ga = a.cpu().numpy()
    g64m = np.array([gb[m] for m in modes])
    g32m = np.array([ga[m] for m in modes])
    scale = np.abs(g64m).max()
    print("  max |g|", scale, "fd-adj64", np.abs(fd - g64m).max(), "fd-adj32",
          np.abs(fd - g32m).max(), flush=True)
    assert np.all(np.abs(fd - g64m) <= 2e-3 * scale + 2e-2 * np.abs(g64m))
    assert np.all(np.abs(fd - g32m) <= 2e-2 * scale + 5e-2 * np.abs(g64m))
    ctx64.close()
    ctx32.close()
