"""Benchmark: IC particles/s at 1024^3 (GRF + 2LPT + TSC displacement + CIC + P(k)).

One step = one realization of the fused FP32 pipeline (white noise from seed
42, BBKS spectrum normalised to sigma8 = 0.81, 2LPT, 1024^3 lattice particles,
CIC deposit, P(k) in 128 bins), box 2000 Mpc/h, all inputs resident in HBM.
The same realization is also run through the plain FP64 path (the reference's
call order, one drop-in call per step) and the fused FP64 path on the same
GPU, and the FP32 P(k)/displacements are compared with the FP64 ones.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--N 1024]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "IC particles/s at 1024^3 (GRF+2LPT+CIC+P(k)), FP32 vs FP64 plain, same GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def b_alg(N: int, s: int) -> float:
    """Algorithmic bytes of one realization: s (28 N^3 + 60 N^2 (N+2))."""
    return s * (28.0 * N ** 3 + 60.0 * N ** 2 * (N + 2))


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "ic_fused_fp32_1024_hbm_traffic.json")


def measured_traffic(stage: str, N: int):
    """HBM bytes per launch of a pipeline stage from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (scaled by the copy-kernel calibration), or
    None when the table does not cover this size and stage."""
    try:
        with open(TRAFFIC_TABLE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    st = t.get("stages", {}).get(stage)
    if t.get("N") != N or not st:
        return None
    return st["traffic_bytes"] / st["launches"]


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(min_seconds: float = 10.0) -> dict:
    """Configuration 1 on the host: the oracle's reference-order FP64 pipeline
    at 64^3 in a 256 Mpc/h box (seed 42, 32 bins), one process (one rank),
    OpenMP only where the reference has it (the kernel loops and the particle
    loops; src/fft.c:139,279, src/mitos.c:844,1072) and serial FFTs (the
    reference's FFTW plans are single-threaded), repeated for min_seconds."""
    from oracle import oracle as O
    O.build()
    n, L = 64, 256.0
    threads = max(1, min(16, os.cpu_count() or 1))
    O.ic_run(n, L, 42, bins=32, threads=threads, fft_threads=1)  # plans, page-in
    reps, t0 = 0, time.perf_counter()
    stages = [0.0] * 7
    while True:
        r = O.ic_run(n, L, 42, bins=32, threads=threads, fft_threads=1)
        stages = [a + b for a, b in zip(stages, r["stage_seconds"])]
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds:
            break
    names = ("grf", "transfer", "2lpt", "displacement_grids", "particles_tsc", "cic", "pk")
    return {"value": reps * n ** 3 / el, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": f"configuration 1: oracle FP64 reference-order pipeline (GRF, 19 FFTs, "
                      f"2LPT, TSC, CIC, P(k) 32 bins) at {n}^3, L={L:g} Mpc/h, seed 42, "
                      f"{reps} runs in {el:.1f} s; 1 process (1 MPI rank), {threads} OpenMP "
                      f"threads in the reference's OpenMP loops, serial FFTs",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "ranks": 1,
            "fft_backend": "the oracle's own FP64 mixed-radix FFT (FFTW is not installed), "
                           "single-threaded as the reference's FFTW plans",
            "stage_seconds_per_run": {k: v / reps for k, v in zip(names, stages)}}


def bench_ensemble(ic, torch, reps: int = 5) -> dict:
    """96 realizations (32 each at 64^3/250, 128^3/500, 256^3/1000 Mpc/h, seeds
    1..96, n_s over [0.94, 0.99], FP32, 32 bins) as one graph launch, against
    the same realizations issued one by one without a graph."""
    cfgs = ic.ensemble_configs()
    log("[bench] ensemble: create + capture")
    ens = ic.Ensemble(cfgs)
    ens.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ens.enqueue()
    ens.run()
    el = (time.perf_counter() - t0) / (reps + 1)
    ens.close()
    npart = sum(c.N ** 3 for c in cfgs)
    balg = sum(b_alg(c.N, 4) for c in cfgs)
    # unbatched: one context per size, realizations issued in sequence
    ctxs = {c.N: ic.IC(c) for c in cfgs}
    stream = torch.cuda.Stream()
    for c in cfgs:
        ctxs[c.N].enqueue(c, stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in cfgs:
        ctxs[c.N].enqueue(c, stream.cuda_stream)
    torch.cuda.synchronize()
    el1 = time.perf_counter() - t0
    for x in ctxs.values():
        x.close()
    log(f"[bench] ensemble: {1e3 * el:.1f} ms per graph launch, unbatched {1e3 * el1:.1f} ms")
    return {"realizations": len(cfgs), "graph_launch_ms": 1e3 * el,
            "realizations_per_s": len(cfgs) / el, "particles_per_s": npart / el,
            "unbatched_ms": 1e3 * el1, "speedup_vs_unbatched": el1 / el,
            "hbm_frac_of_B_alg": balg / el / 1e9 / PEAK_HBM_GBS,
            "workload": "32 x 64^3/250 + 32 x 128^3/500 + 32 x 256^3/1000 Mpc/h, seeds 1..96, "
                        "n_s 0.94..0.99, FP32, 2LPT, CIC, P(k) 32 bins"}


def bench_adjoint(ic, torch, reps: int = 3) -> dict:
    """Configuration 5: dL/d(white noise) of L = sum_b (P_b - P_target,b)^2 / b
    at 256^3 in a 500 Mpc/h box (seed 7, target seed 8, 32 bins, FP32): one
    forward (P(k)) and one backward per step, timed after a warm-up, and the
    gradient's bit-reproducibility between two steps."""
    from mitos_amd import adjoint
    N, L, bins = 256, 500.0, 32
    cfg = ic.ICConfig(N=N, boxlen=L, seed=7, bins=bins, precision=ic.MITOS_FP32)
    ctx = ic.IC(cfg)
    tgt = ctx.noise_forward(cfg, ic.white_noise(N, 8, ic.MITOS_FP32))["P"]
    noise = ic.white_noise(N, 7, ic.MITOS_FP32)

    def step():
        leaf = noise.detach().requires_grad_(True)
        adjoint.pk_loss(adjoint.noise_power_spectrum(leaf, ctx, cfg), tgt).backward()
        return leaf.grad

    g0 = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g = step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    same = bool(torch.equal(g0, g))
    ctx.close()
    log(f"[bench] adjoint 256^3 fp32: {1e3 * el:.1f} ms per forward + backward")
    return {"workload": "configuration 5: 256^3, 500 Mpc/h, seed 7 (target seed 8), 32 bins, "
                        "FP32, P(k) loss gradient w.r.t. the white noise",
            "ms_per_gradient": 1e3 * el, "bit_reproducible": same}


def pmc_pass(ic, torch, N, L, bins):
    """Dispatches profiled by the counter runs: realization 2 is the measured one."""
    cfg = ic.ICConfig(N=N, boxlen=L, seed=42, bins=bins, precision=ic.MITOS_FP32)
    ctx = ic.IC(cfg)
    stream = torch.cuda.Stream()
    for i in range(2):
        if i == 1:
            ctx.profile(True)  # events only: the dispatch sequence is unchanged
        ctx.enqueue(cfg, stream.cuda_stream)
        torch.cuda.synchronize()
        log(f"[pmc] realization {i + 1} done")
    log("[pmc] stage order of realization 2: " +
        ",".join(name for name, _, _ in ctx.profile_entries()))
    ctx.close()
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)  # copy calibration kernel: reads 1 GiB, writes 1 GiB
    torch.cuda.synchronize()
    log("[pmc] copy calibration done (1 GiB read, 1 GiB written)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--boxlen", type=float, default=2000.0)
    ap.add_argument("--bins", type=int, default=128)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparisons")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 96-realization ensemble (one graph launch)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the configuration-5 gradient timing")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="one warm-up and one measured FP32 realization, then a plain copy "
                         "of 1 GiB (counter calibration); for rocprofv3 --pmc runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU (requested --gpus {a.gpus}); this build has no "
              "multi-GPU path: the slab/pencil sharding is a design only.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import mitos_amd.ic as ic

    N, L, bins = a.N, a.boxlen, a.bins
    if a.pmc_pass:
        pmc_pass(ic, torch, N, L, bins)
        return
    out = {"metric": METRIC, "unit": "particles/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    out["config"] = {"workload": f"Mitos IC realization {N}^3, box {L:g} Mpc/h, seed 42, BBKS "
                                 f"Gamma=0.21 n_s=0.965 sigma8=0.81, 2LPT, TSC displacement, "
                                 f"CIC, P(k) {bins} bins; FP32 fields, FP64 P(k) sums",
                     "N": N, "boxlen": L, "seed": 42, "bins": bins, "precision": "fp32",
                     "pipeline": "fused"}

    # ---- FP32 fused: the timed path -------------------------------------------------------
    cfg = ic.ICConfig(N=N, boxlen=L, seed=42, bins=bins, precision=ic.MITOS_FP32)
    log(f"[bench] fp32 fused {N}^3: create")
    ctx = ic.IC(cfg)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    for i in range(a.warmup):
        ctx.enqueue(cfg, sp)
    torch.cuda.synchronize()
    log(f"[bench] warmup done; timing {a.steps} steps")
    t0 = time.perf_counter()
    for i in range(a.steps):
        ctx.enqueue(cfg, sp)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = 1e3 * el / a.steps
    out["value"] = N ** 3 * a.steps / el
    out["ms_per_step"] = ms
    out["fft_schedules"] = {k: {"variant": v, "ms_per_variant": ms}
                            for k, (v, ms) in ctx.schedules().items()}
    # one more run with the per-kernel profile and the P(k) result
    ctx.profile(True)
    r32 = ctx.run(cfg)
    prof = ctx.profile_entries()
    ctx.profile(False)
    tot = {}
    for name, kms, by in prof:
        t = tot.setdefault(name, [0.0, 0.0, 0])
        t[0] += kms
        t[1] += by
        t[2] += 1
    dom = max(tot.items(), key=lambda kv: kv[1][0])
    dname, (dms, dby, dn) = dom
    achieved = dby / (dms * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                       "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                       "traffic": measured_traffic(dname, N),
                       "kernel": dname, "launches": dn, "kernel_ms": dms / dn,
                       "algorithmic_bytes_per_launch": dby / dn}
    out["kernels"] = {k: {"ms": v[0], "GB_per_s": v[1] / (v[0] * 1e-3) / 1e9 if v[0] else None,
                          "launches": v[2]} for k, v in tot.items()}
    pipe_frac = b_alg(N, 4) / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS
    out["pipeline_roofline"] = {"algorithmic_bytes": b_alg(N, 4), "achieved_GBs":
                                b_alg(N, 4) / (ms * 1e-3) / 1e9, "frac": pipe_frac,
                                "formula": "s*(28 N^3 + 60 N^2 (N+2)), s=4"}
    log(f"[bench] fp32: {ms:.1f} ms/step, {out['value']:.3e} particles/s; dominant {dname}")
    psi32 = None
    if not a.no_fp64:
        psi32 = ctx.displacements()
    # plain FP32 3D FFTs of this size against B_fft = s (N^3 + 5 N^2 (N+2))
    t_r2c, t_c2r = ctx.fft_timing(5)
    bf = 4.0 * (N ** 3 + 5.0 * N ** 2 * (N + 2))
    out["fft_roofline"] = {"B_fft": bf, "r2c_ms": t_r2c, "c2r_ms": t_c2r,
                           "r2c_frac": bf / (t_r2c * 1e-3) / 1e9 / PEAK_HBM_GBS,
                           "c2r_frac": bf / (t_c2r * 1e-3) / 1e9 / PEAK_HBM_GBS,
                           "formula": "s*(N^3 + 5 N^2 (N+2)), s=4"}
    log(f"[bench] fft {N}^3 fp32: r2c {t_r2c:.2f} ms, c2r {t_c2r:.2f} ms")
    ctx.close()
    del ctx

    if not a.no_ensemble:
        out["ensemble"] = bench_ensemble(ic, torch)
    if not a.no_adjoint:
        out["adjoint"] = bench_adjoint(ic, torch)

    if not a.no_fp64:
        # ---- FP64 fused and plain on the same realization ----------------------------------
        cfg64 = ic.ICConfig(N=N, boxlen=L, seed=42, bins=bins, precision=ic.MITOS_FP64)
        log("[bench] fp64 fused")
        c64 = ic.IC(cfg64)
        c64.run(cfg64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r64 = c64.run(cfg64)
        el64 = time.perf_counter() - t0
        psi64 = c64.displacements()
        c64.close()
        del c64
        d = (psi32.double() - psi64)
        rms = float(torch.sqrt((d * d).mean()) / torch.sqrt((psi64 * psi64).mean()))
        mx = float(d.abs().max() / psi64.abs().max())
        del d, psi32, psi64
        torch.cuda.empty_cache()
        ok = r64["obs"] > 1
        pk_err = np.abs(r32["P"][ok] - r64["P"][ok]) / np.abs(r64["P"][ok])
        log("[bench] fp64 plain (reference call order): warm-up, then timed")
        ic.run_reference_order(cfg64)  # library workspaces, tables, twiddles
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rp = ic.run_reference_order(cfg64)
        elp = time.perf_counter() - t0
        okp = rp["obs"] > 1
        pk_plain = np.abs(rp["P"][okp] - r64["P"][okp]) / np.abs(r64["P"][okp])
        stage_names = ("noise", "transfer", "2lpt", "displacement", "particles", "cic", "pk",
                       "total")
        out["fp64_plain"] = {"particles_per_s": N ** 3 / elp, "ms": 1e3 * elp,
                             "stage_ms": dict(zip(stage_names, rp["stage_ms"])),
                             "pipeline": "reference call order, drop-in calls, 19 FFTs; grids "
                                         "allocated and freed by the calls as in the reference; "
                                         "second run timed"}
        out["fp64_fused"] = {"particles_per_s": N ** 3 / el64, "ms": 1e3 * el64,
                             "stage_ms": dict(zip(stage_names, r64["stage_ms"]))}
        out["speedup_fp32_vs_fp64_plain"] = out["value"] / out["fp64_plain"]["particles_per_s"]
        out["speedup_fp32_vs_fp64_fused"] = out["value"] / out["fp64_fused"]["particles_per_s"]
        out["precision"] = {
            "pk_max_rel_err_fp32_vs_fp64": float(pk_err.max()),
            "pk_rms_rel_err_fp32_vs_fp64": float(np.sqrt((pk_err ** 2).mean())),
            "displacement_rms_rel_err_fp32_vs_fp64": rms,
            "displacement_max_err_over_max_fp32_vs_fp64": mx,
            "pk_max_rel_diff_fp64_fused_vs_plain": float(pk_plain.max()),
            "bins_compared": int(ok.sum())}
        log(f"[bench] fp64 fused {1e3 * el64:.1f} ms, plain {1e3 * elp:.1f} ms; "
            f"P(k) fp32 err {pk_err.max():.2e}")

    if not a.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
