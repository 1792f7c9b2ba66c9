"""Write GGFT golden vectors from the reference solver built from its sources.

Runs the reference ggft/src/ggft_solver.cpp + vendored Eigen, compiled by
oracle/Makefile (oracle/_ref/ggft_ref at N = 23 as the sources stand;
oracle/_ref/ggft_ref_n16 / _n21 from copies that change only the compile-time
N of ggft_solver.h:33), at tolerance 1e-13 on seeded structures and stores,
per case, the structure description (inputs) and the reference G (expected
output).  The permittivity cube is recovered by voxelizing the description at
N without normalisation, so the fixture holds inputs and outputs only.

    python tests/golden/make_ggft_golden.py            # N = 23 (ggft_n23.npz)
    python tests/golden/make_ggft_golden.py 16 21      # ggft_n16.npz, ggft_n21.npz

Even N takes the reference's separate even-N source and gradient stencils
(ggft_solver.cpp:178-273), so the N = 16 file covers every right-hand side
at even N.  Random layered/box cubes mostly leave the cells around the source
homogeneous, where the plus-side and minus-side terms of the gradient
stencils coincide and a mirrored or axis-swapped stencil gives the same G, so
every file also holds "centre" cases (centre_cases): an interface through
the source along z, along x and along y, and corner boxes that make all eight
(even N) or all three (odd N) stencil cells differ, each for WVTZ, WVTX and
WVTY, and the asymmetric cube for GFT.
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

def out_path(N):
    return os.path.join(ROOT, "tests", "golden", f"ggft_n{N}.npz")


def run_ref(cubes, code, N, tol=1e-13):
    exe = os.path.join(ROOT, "oracle", "_ref", "ggft_ref" if N == 23 else f"ggft_ref_n{N}")
    inp = struct.pack("iid", len(cubes), code, tol) + b"".join(np.asarray(c, np.float64).tobytes() for c in cubes)
    res = subprocess.run([exe], input=inp, capture_output=True, check=True)
    a = np.frombuffer(res.stdout, dtype=np.float64).reshape(len(cubes), 6 * N * N + 1)
    return a[:, :-1].copy(), a[:, -1].copy()


def centre_cases(N, seed):
    """Descriptions whose permittivity changes across the source: the plane
    between cells H-1 and H (even N, at 0) or between H and H+1 (odd N, at
    1/N) carries an interface normal to z, x or y, or the corner of two boxes."""
    rng = np.random.default_rng(seed)
    b = 0.0 if N % 2 == 0 else 1.0 / N
    descs = []
    for kind in ("z", "x", "y", "corner"):
        e = rng.uniform(1.0, 8.0, 4)
        L = np.full((15, 2), structures.PAD, np.float32)
        C = np.full((15, 7), structures.PAD, np.float32)
        L[0] = (-1.5, e[0])
        if kind == "z":
            L[1] = (b, e[1])
        elif kind == "x":
            C[0] = (b, -1.5, -1.5, 1.5, 1.5, 1.5, e[1])
        elif kind == "y":
            C[0] = (-1.5, b, -1.5, 1.5, 1.5, 1.5, e[1])
        else:
            L[1] = (b, e[1])
            C[0] = (b, b, -1.5, 1.5, 1.5, 1.5, e[2])
            C[1] = (-1.5, b, -1.5, b, 1.5, b, e[3])
        descs.append((L, C))
    return descs


def write(N):
    if N == 23:
        # 3 Poisson (GFT) cases, 3 WVTZ, 1 WVTX, 1 WVTY; plus the homogeneous cube.
        codes, n_homog, seed = [0, 0, 0, 1, 1, 1, 2, 3, 0, 1], 2, 2024
    else:
        # two layered/box cubes per code, then the homogeneous cube per code
        codes, n_homog, seed = [0, 0, 1, 1, 2, 2, 3, 3, 0, 1, 2, 3], 4, 2024 + N
    B = len(codes)
    layers, cuboids, _ = structures.layered_boxes(B, N, seed=seed, max_boxes=2)
    # last cases: homogeneous eps = 1 (no boxes, one layer)
    for b in range(B - n_homog, B):
        layers[b] = structures.PAD
        cuboids[b] = structures.PAD
        layers[b, 0] = (-1.5, 1.0)
    # centre cases: each interface / corner cube for WVTZ, WVTX and WVTY, and
    # the corner cube for GFT
    cen = centre_cases(N, seed + 7)
    extra = [(d, c) for d in cen for c in (1, 2, 3)] + [(cen[3], 0)]
    layers = np.concatenate([layers, np.stack([d[0] for d, _ in extra])])
    cuboids = np.concatenate([cuboids, np.stack([d[1] for d, _ in extra])])
    codes = codes + [c for _, c in extra]
    B = len(codes)
    cubes, _ = O.voxelize(layers, cuboids, N, normalise=False)
    G = np.zeros((B, 6 * N * N))
    secs = np.zeros(B)
    for code in sorted(set(codes)):
        idx = [i for i, c in enumerate(codes) if c == code]
        g, s = run_ref([cubes[i] for i in idx], code, N)
        G[idx] = g
        secs[idx] = s
    out = out_path(N)
    np.savez_compressed(out, layers=layers, cuboids=cuboids, codes=np.array(codes, np.int32), G=G, N=N,
                        ref_seconds=secs)
    print("wrote", out, "ref seconds per solve", secs)


def main():
    O.build()
    for N in [int(a) for a in sys.argv[1:]] or [23]:
        write(N)


if __name__ == "__main__":
    main()
