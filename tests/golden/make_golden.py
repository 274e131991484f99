"""Write the golden vectors under tests/golden/ (run once, where the reference
tree and a C compiler are present):

  python tests/golden/make_golden.py

xoshiro.npz   uint64 draws of the reference's own include/random_xorshift.h
              (xoshiro256** seeded through splitmix64), produced by
              oracle/_ref/xoshiro_ref, which oracle/Makefile (target ref)
              compiles from the reference tree in place:
                first<seed>       first 4096 draws of seeds 42, 100, 7, 8
                seeds1to96        first 16 draws of seeds 1..96, shape (96, 16)
                off_skips, off_draws
                                  8 draws of seed 42 after skipping 4(m-1) draws,
                                  i.e. the four draws of mode m and the next mode,
                                  for sample modes m up to the last mode at 1024^3
normals.npz   Box-Muller normals of src/random.c:34-49 (cosine branch, two
              draws per normal) computed here in Python from the seed-42 draws
              with the platform's libm, 2048 values.
"""
from __future__ import annotations

import math
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "xoshiro_ref")


def ref_draws(seed: int, skip: int, n: int) -> np.ndarray:
    out = subprocess.run([REF, str(seed), str(skip), str(n)], check=True, capture_output=True,
                         text=True).stdout.split()
    return np.array([int(v) for v in out], dtype=np.uint64)


def modes(N: int) -> int:
    """Non-DC modes of an N^3 half grid (each consumes 4 draws)."""
    return N * N * (N // 2 + 1) - 1


def box_muller(a: int, b: int) -> float:
    # src/random.c:38-46: u = ((double)A + 0.5) / RMax, RMax = (double)UINT64_MAX + 1
    rmax = float(2 ** 64 - 1) + 1.0
    u = (float(a) + 0.5) / rmax
    v = (float(b) + 0.5) / rmax
    return math.sqrt(-2.0 * math.log(u)) * math.cos(2.0 * math.pi * v)


def main() -> None:
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {}
    for seed in (42, 100, 7, 8):
        out[f"first{seed}"] = ref_draws(seed, 0, 4096)
    out["seeds1to96"] = np.stack([ref_draws(s, 0, 16) for s in range(1, 97)])
    ms = [1, 2, 1000, 123457, modes(64), modes(128), modes(256), 30_000_001, modes(512),
          modes(1024)]
    skips = np.array([4 * (m - 1) for m in ms], dtype=np.uint64)
    out["off_skips"] = skips
    out["off_draws"] = np.stack([ref_draws(42, int(s), 8) for s in skips])
    np.savez(os.path.join(HERE, "xoshiro.npz"), **out)
    d = out["first42"]
    z = np.array([box_muller(int(d[2 * i]), int(d[2 * i + 1])) for i in range(2048)])
    np.savez(os.path.join(HERE, "normals.npz"), seed42=z)
    print("wrote", os.path.join(HERE, "xoshiro.npz"), "and normals.npz", file=sys.stderr)


if __name__ == "__main__":
    main()
