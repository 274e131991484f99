"""torch.autograd front end of the hand-written adjoint in libmitos_hip.so.

    P = noise_power_spectrum(noise, ic, cfg)   # (bins,) float64, CPU
    delta = noise_density(noise, ic, cfg)      # (N, N, N) CUDA, context real type

`noise` is unit-variance complex white noise of shape (N, N, N/2+1) on the
GPU (``mitos_amd.ic.white_noise`` gives the reference seed stream's).  The
forward runs the IC pipeline (GRF amplitude, 2LPT, TSC displacement, CIC,
P(k)); the backward is mitos_ic_noise_backward: no graph is recorded through
the pipeline, and the Hessian grids are recomputed, not stored.
"""
from __future__ import annotations

import numpy as np
import torch


class NoisePowerSpectrum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, noise, ic, cfg):
        out = ic.noise_forward(cfg, noise.detach())
        ctx.ic, ctx.cfg = ic, cfg
        ctx.save_for_backward(noise.detach())
        return torch.from_numpy(out["P"].copy())

    @staticmethod
    def backward(ctx, gP):
        noise, = ctx.saved_tensors
        g = gP.detach().to("cpu", torch.float64).numpy()
        return ctx.ic.noise_backward(ctx.cfg, noise, dL_dP=g), None, None


class NoiseDensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, noise, ic, cfg):
        ic.noise_forward(cfg, noise.detach())
        ctx.ic, ctx.cfg = ic, cfg
        ctx.save_for_backward(noise.detach())
        return ic.density().reshape(cfg.N, cfg.N, cfg.N)

    @staticmethod
    def backward(ctx, gd):
        noise, = ctx.saved_tensors
        return ctx.ic.noise_backward(ctx.cfg, noise, dL_ddelta=gd), None, None


def noise_power_spectrum(noise, ic, cfg):
    return NoisePowerSpectrum.apply(noise, ic, cfg)


def noise_density(noise, ic, cfg):
    return NoiseDensity.apply(noise, ic, cfg)


def pk_loss(P, P_target):
    """sum_b w_b (P_b - P_target,b)^2 with w_b = 1/b (b = 1-based bin), over
    bins where both spectra are defined."""
    P_target = torch.as_tensor(np.asarray(P_target), dtype=torch.float64)
    w = 1.0 / torch.arange(1, P.numel() + 1, dtype=torch.float64)
    ok = ~(torch.isnan(P.detach()) | torch.isnan(P_target))
    return (w[ok] * (P[ok] - P_target[ok]) ** 2).sum()
