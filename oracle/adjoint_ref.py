"""FP64 reference gradient of the IC pipeline for the adjoint tests.

TEST INFRASTRUCTURE ONLY (the checker of the hand-written adjoint): the
forward that mitos_ic_noise_forward runs, restated with torch operations on
the CPU in float64, differentiated by torch.autograd.  Nothing here is
hand-differentiated, so it is an independent gradient.

The forward (stencil form, positions kept in float64):
  K      = noise * sqrt(L^3/2) * sqrt(A k^n_s) T_BBKS(k) * (-1/k^2), Hermitian
           planes as src/grf.c:78-175
  phi_ij = c2r(-k_i k_j K) / L^3; delta2 = sum of minors (src/2lpt.c:111-112)
  Phi    = W_TSC^-1 (f1 K + f2 (L^3/N^3) ip(k) r2c(delta2)), ip = -1/k^2 (1 at DC)
  psi_a  = c2r(i k_a S.. Phi) / L^3 with the lattice stencils S (3/4+cos/4)
  moves  : x -= psi_x(q); y -= TSC_x psi_y; z -= TSC_xy psi_z (src/mitos.c:827-863)
  CIC    : weights of analyse/src/mitos_read.c:285-320, delta = rho - 1
  P(k)   : src/calc_powerspec.c:32-90 with the CIC window (order 2) and the
           r2c normalisation L^3/N^3.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def bbks(k, gamma):
    q = k / gamma
    return (torch.log(1 + 2.34 * q) / (2.34 * q) *
            (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25)


def _freq(N):
    f = np.arange(N)
    return np.where(f > N // 2, f - N, f)  # index N/2 -> +N/2 (src/fft.c:31-33)


def _sinc(x):
    return np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))


def pk_bins(N, L, bins):
    """bin index of every half-grid mode (-1 for DC), float32 u as the reference"""
    dk = 2 * math.pi / L
    fx = _freq(N)
    kx = (fx * dk)[:, None, None]
    ky = (fx * dk)[None, :, None]
    kz = (np.arange(N // 2 + 1) * dk)[None, None, :]
    k = np.sqrt(kx * kx + ky * ky + kz * kz)
    lmin, lmax = math.log(dk), math.log(math.sqrt(3) * dk * N / 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = ((np.log(k) - lmin) / (lmax - lmin)).astype(np.float32)
        b = np.floor(np.float32(bins - 1) * u)
    b = np.where(k > 0, b, -1).astype(np.int64)
    return b


class Forward:
    def __init__(self, N, L, bins, amplitude, n_s=0.965, gamma=0.21, f1=1.0, f2=-3.0 / 7.0,
                 run2lpt=True):
        self.N, self.L, self.bins = N, L, bins
        self.f1, self.f2 = f1, (f2 if run2lpt else 0.0)
        self.run2lpt = run2lpt
        dk = 2 * math.pi / L
        fx = _freq(N)
        H = N // 2 + 1
        kx = torch.tensor(fx * dk)[:, None, None].expand(N, N, H)
        ky = torch.tensor(fx * dk)[None, :, None].expand(N, N, H)
        kz = torch.tensor(np.arange(H) * dk)[None, None, :].expand(N, N, H)
        self.k = (kx, ky, kz)
        k2 = kx * kx + ky * ky + kz * kz
        k = torch.sqrt(k2)
        ks = torch.where(k > 0, k, torch.ones_like(k))
        amp = torch.sqrt(amplitude * ks ** n_s) * bbks(ks, gamma) * (-1.0 / ks / ks)
        self.amp = torch.where(k > 0, amp, torch.zeros_like(amp)) * math.sqrt(L ** 3 / 2)
        self.ip = torch.where(k2 > 0, -1.0 / torch.where(k2 > 0, k2, torch.ones_like(k2)),
                              torch.ones_like(k2))
        w3 = torch.tensor(1.0 / _sinc(0.5 * fx * dk * L / N) ** 3)
        w3z = torch.tensor(1.0 / _sinc(0.5 * np.arange(H) * dk * L / N) ** 3)
        self.win3 = w3[:, None, None] * w3[None, :, None] * w3z[None, None, :]
        S = torch.tensor(0.75 + 0.25 * np.cos(2 * np.pi * np.arange(N) / N))
        self.Sx, self.Sy, self.Sz = S[:, None, None], S[None, :, None], S[None, None, :H]
        w2 = 1.0 / _sinc(0.5 * fx * dk * L / N) ** 2
        w2z = 1.0 / _sinc(0.5 * np.arange(H) * dk * L / N) ** 2
        self.win2 = torch.tensor(w2[:, None, None] * w2[None, :, None] * w2z[None, None, :])
        b = pk_bins(N, L, bins)
        self.bin = torch.tensor(b)
        m = np.full(b.shape, 2.0)
        m[:, :, 0] = 1
        m[:, :, N // 2] = 1
        self.mult = torch.tensor(m)
        sel = b >= 0
        self.obs = np.bincount(b[sel], weights=m[sel], minlength=bins)
        # Hermitian planes: lower points and their mirrors (src/grf.c:118-137)
        low, mir, selfc = [], [], []
        for z in (0, N // 2):
            for x in range(N // 2 + 1):
                for y in range(N):
                    if (x in (0, N // 2)) and y > N // 2:
                        continue
                    ix, iy = (-x) % N, (-y) % N
                    if ix == x and iy == y:
                        selfc.append((x, y, z))
                    else:
                        low.append((x, y, z))
                        mir.append((ix, iy, z))
        self.low = tuple(torch.tensor(v) for v in zip(*low))
        self.mir = tuple(torch.tensor(v) for v in zip(*mir))
        self.selfc = tuple(torch.tensor(v) for v in zip(*selfc))

    def c2r(self, X):
        N = self.N
        return torch.fft.irfftn(X, s=(N, N, N)) * N ** 3

    def r2c(self, x):
        return torch.fft.rfftn(x)

    def hermitian(self, K):
        K = K.clone()
        K[self.low] = torch.conj(K[self.mir])
        K[self.selfc] = torch.complex(K[self.selfc].real, torch.zeros_like(K[self.selfc].real))
        return K

    def density(self, noise):
        N, L = self.N, self.L
        V = L ** 3
        kx, ky, kz = self.k
        K = self.hermitian(noise * self.amp)
        Phi = self.f1 * K
        if self.run2lpt:
            h = {}
            for name, fac in (("xx", -kx * kx), ("yy", -ky * ky), ("zz", -kz * kz),
                              ("xy", -kx * ky), ("yz", -ky * kz), ("xz", -kx * kz)):
                h[name] = self.c2r(fac * K) / V
            d2 = (h["zz"] * h["yy"] + h["zz"] * h["xx"] + h["yy"] * h["xx"]
                  - h["xy"] ** 2 - h["yz"] ** 2 - h["xz"] ** 2)
            Phi = Phi + self.f2 * (V / N ** 3) * self.ip * self.r2c(d2)
        Phi = Phi * self.win3
        psi_x = self.c2r(1j * kx * self.Sx * self.Sy * self.Sz * Phi) / V
        psi_y = self.c2r(1j * ky * self.Sy * self.Sz * Phi) / V
        psi_z = self.c2r(1j * kz * self.Sz * Phi) / V
        # moves of the lattice particles
        sp = float(np.float32(L) / np.float32(N))
        q = torch.arange(N)
        qx, qy, qz = torch.meshgrid(q, q, q, indexing="ij")
        x = qx * sp - psi_x
        y = (qy * sp).to(torch.float64)
        z = (qz * sp).to(torch.float64)
        sN = N / L

        def taps(X):
            c = torch.floor(X + 0.5)
            u = X - c
            c = c.long()
            w = ((0.5 * (0.5 - u) ** 2), 0.75 - u * u, 0.5 * (0.5 + u) ** 2)
            return [((c + t - 1) % N) for t in range(3)], w

        ix, wx = taps(x * sN)
        y = y - sum(wx[a] * psi_y[ix[a], qy, qz] for a in range(3))
        iy, wy = taps(y * sN)
        z = z - sum(wx[a] * wy[b] * psi_z[ix[a], iy[b], qz] for a in range(3) for b in range(3))
        # CIC
        rho = torch.zeros(N ** 3, dtype=torch.float64)
        cells, ws = [], []
        for v in (x, y, z):
            X = torch.remainder(v, L) * sN
            i = torch.floor(X)
            d = X - i
            i = i.long() % N
            cells.append((i, (i + 1) % N))
            ws.append((1 - d, d))
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    idx = (cells[0][a] * N + cells[1][b]) * N + cells[2][c]
                    rho = rho.index_add(0, idx.reshape(-1),
                                        (ws[0][a] * ws[1][b] * ws[2][c]).reshape(-1))
        return rho.reshape(N, N, N) - 1.0

    def power(self, delta):
        N, L = self.N, self.L
        Kd = self.r2c(delta) * (L ** 3 / N ** 3) * self.win2
        pw = self.mult * (Kd.real ** 2 + Kd.imag ** 2)
        sel = self.bin >= 0
        sums = torch.zeros(self.bins, dtype=torch.float64).index_add(0, self.bin[sel], pw[sel])
        obs = torch.tensor(self.obs)
        return sums / obs / L ** 3


def pk_loss(P, P_target):
    w = 1.0 / torch.arange(1, P.numel() + 1, dtype=torch.float64)
    T = torch.as_tensor(np.asarray(P_target), dtype=torch.float64)
    ok = ~(torch.isnan(P.detach()) | torch.isnan(T))
    return (w[ok] * (P[ok] - T[ok]) ** 2).sum()


def pk_gradient(fw: Forward, noise: np.ndarray, P_target) -> tuple:
    """(loss, dL/dnoise as complex d/dRe + i d/dIm) for the P(k) loss"""
    nz = torch.tensor(noise, dtype=torch.complex128, requires_grad=True)
    L = pk_loss(fw.power(fw.density(nz)), P_target)
    L.backward()
    return float(L.detach()), nz.grad.numpy()


def density_gradient(fw: Forward, noise: np.ndarray, r: np.ndarray) -> tuple:
    """(loss, gradient) for L = sum r delta + delta^2 / 2"""
    nz = torch.tensor(noise, dtype=torch.complex128, requires_grad=True)
    d = fw.density(nz)
    rt = torch.tensor(r)
    L = (d * rt).sum() + 0.5 * (d * d).sum()
    L.backward()
    return float(L.detach()), nz.grad.numpy()
