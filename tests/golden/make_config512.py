"""Write tests/golden/config512.npz: the CPU oracle's FP64 reference-order
pipeline at 512^3 in a 1000 Mpc/h box (seed 42, BBKS Gamma = 0.21, n_s =
0.965, sigma8 = 0.81, 2LPT, 128 P(k) bins), reduced to what a GPU test can
compare without re-running the oracle:

  k, P, obs              the 128 bins of calc_cross_powerspec
  psi_norm[3]            L2 norms of the three displacement grids (unpadded)
  psi_idx, psi_val       4096 sampled values of each displacement grid
                         (flat index into the unpadded N^3 grid, per component)
  dens_norm, dens_mean   L2 norm and mean of the CIC overdensity
  dens_idx, dens_val     4096 sampled overdensity values
  part_idx, X, Y, Z      4096 sampled particle positions (float, after the
                         three sequential TSC moves and the wrap)
  pos_checksum           float64 sums of X, Y, Z over all particles

Run once in this container (about two minutes with 8 threads):

  python tests/golden/make_config512.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

N, L, SEED, BINS = 512, 1000.0, 42, 128
NSAMP = 4096


def main() -> None:
    from oracle import oracle as O
    O.build()
    t0 = time.time()
    ref = O.ic_run(N, L, SEED, bins=BINS, want_psi=True, want_density=True,
                   threads=os.cpu_count() or 1)
    print(f"oracle 512^3 run: {time.time() - t0:.1f} s", file=sys.stderr)
    rng = np.random.default_rng(20512)
    n = N ** 3
    psi = ref["psi"]
    out = {"k": ref["k"], "P": ref["P"], "obs": ref["obs"]}
    out["psi_norm"] = np.array([np.linalg.norm(psi[i, :, :, :N]) for i in range(3)])
    psi_idx = rng.integers(0, n, size=(3, NSAMP))
    psi_val = np.empty((3, NSAMP))
    for i in range(3):
        xi, r = np.divmod(psi_idx[i], N * N)
        yi, zi = np.divmod(r, N)
        psi_val[i] = psi[i, xi, yi, zi]
    out["psi_idx"] = psi_idx
    out["psi_val"] = psi_val
    del psi, ref["psi"]
    dens = ref["density"].reshape(-1)
    out["dens_norm"] = np.array(np.linalg.norm(dens))
    out["dens_mean"] = np.array(dens.mean())
    out["dens_idx"] = rng.integers(0, n, size=NSAMP)
    out["dens_val"] = dens[out["dens_idx"]]
    pidx = rng.integers(0, n, size=NSAMP)
    out["part_idx"] = pidx
    for c in "XYZ":
        out[c] = ref[c][pidx]
    out["pos_checksum"] = np.array([ref[c].astype(np.float64).sum() for c in "XYZ"])
    path = os.path.join(HERE, "config512.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, f"({os.path.getsize(path)} bytes)", file=sys.stderr)


if __name__ == "__main__":
    main()
