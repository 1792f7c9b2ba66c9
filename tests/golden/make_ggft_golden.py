"""Write GGFT golden vectors from the reference solver built from its sources.

Runs oracle/_ref/ggft_ref (the reference ggft/src/ggft_solver.cpp + vendored
Eigen, compiled by oracle/Makefile; N = 23 is fixed by ggft_solver.h:33) at
tolerance 1e-13 on seeded structures and stores, per case, the structure
description (inputs) and the reference G (expected output).  The permittivity
cube is recovered by voxelizing the description at N = 23 without
normalisation, so the fixture holds inputs and outputs only.

    python tests/golden/make_ggft_golden.py
"""
from __future__ import annotations

import os
import struct
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from deeprwcap_amd import structures  # noqa: E402
from oracle import oracle as O  # noqa: E402

N = 23
OUT = os.path.join(ROOT, "tests", "golden", "ggft_n23.npz")


def run_ref(cubes, code, tol=1e-13):
    exe = os.path.join(ROOT, "oracle", "_ref", "ggft_ref")
    inp = struct.pack("iid", len(cubes), code, tol) + b"".join(np.asarray(c, np.float64).tobytes() for c in cubes)
    res = subprocess.run([exe], input=inp, capture_output=True, check=True)
    a = np.frombuffer(res.stdout, dtype=np.float64).reshape(len(cubes), 6 * N * N + 1)
    return a[:, :-1].copy(), a[:, -1].copy()


def main():
    O.build()
    # 3 Poisson (GFT) cases, 3 WVTZ, 1 WVTX, 1 WVTY; plus the homogeneous cube.
    codes = [0, 0, 0, 1, 1, 1, 2, 3, 0, 1]
    B = len(codes)
    layers, cuboids, _ = structures.layered_boxes(B, N, seed=2024, max_boxes=2)
    # last two cases: homogeneous eps = 1 (no boxes, one layer)
    for b in (B - 2, B - 1):
        layers[b] = structures.PAD
        cuboids[b] = structures.PAD
        layers[b, 0] = (-1.5, 1.0)
    cubes, _ = O.voxelize(layers, cuboids, N, normalise=False)
    G = np.zeros((B, 6 * N * N))
    secs = np.zeros(B)
    for code in sorted(set(codes)):
        idx = [i for i, c in enumerate(codes) if c == code]
        g, s = run_ref([cubes[i] for i in idx], code)
        G[idx] = g
        secs[idx] = s
    np.savez_compressed(OUT, layers=layers, cuboids=cuboids, codes=np.array(codes, np.int32), G=G, N=N,
                        ref_seconds=secs)
    print("wrote", OUT, "ref seconds per solve", secs)


if __name__ == "__main__":
    main()
