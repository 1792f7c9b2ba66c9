"""GGFT dataset files (poisson.bin / gradient.bin, the reference's training
input; ggft/src/ggft_solver.cpp:35-48, ggft_solver.h:180-191, main.cpp:59-78)
written by deeprwcap_amd.training on the CPU:

* the reference solver, compiled from its own sources (oracle/_ref/ggft_ref)
  and given a file name, writes the header and a record per solve itself; for
  the same permittivities and structure blocks and the G it computed, our
  writer produces the same bytes, and our reader returns the inputs;
* the C++ facade's DNNSolver::readBinaryFile (structureSize = 7 x 16) reads
  our file as the values written (as float32);
* write/append/read round trip, and the structure blocks of a description
  (the generator's centre / half-extent form, first match wins).
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from deeprwcap_amd import structures as S
from deeprwcap_amd import training as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "ggft_ref")
N = 23


def _descriptions(b, seed):
    L, C, _ = S.layered_boxes(b, N, seed=seed, max_boxes=2)
    return L, C


def test_blocks_of_a_description():
    L = np.full((1, 15, 2), S.PAD, np.float32)
    C = np.full((1, 15, 7), S.PAD, np.float32)
    L[0, :2] = [(-1.5, 2.0), (0.2, 4.5)]
    C[0, 0] = (-0.5, -0.25, -1.5, 0.5, 0.75, 1.5, 7.0)
    blk = T.description_blocks(L, C)
    assert blk.shape == (1, 16, 7)
    np.testing.assert_allclose(blk[0, 0], [0, 0.25, 0, 0.5, 0.5, 1.5, 7.0])
    np.testing.assert_allclose(blk[0, 1], [0, 0, -0.9, 2, 2, 1.1, 2.0], rtol=1e-6)   # z in [-2, 0.2]
    np.testing.assert_allclose(blk[0, 2], [0, 0, 1.1, 2, 2, 0.9, 4.5], rtol=1e-6)    # z in [0.2, 2]
    assert (blk[0, 3:] == 0).all()
    # voxel centres take the first block holding them (|x - c| < r), as the
    # generator fills DIEL (dielectric_generator.h:110-123)
    from oracle import oracle as O
    cube, _ = O.voxelize(L, C, N, normalise=False)
    t = (2 * np.arange(N) + 1) / N - 1
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    val = np.zeros((N, N, N))
    done = np.zeros((N, N, N), bool)
    for b in blk[0]:
        if b[6] == 0:
            continue
        inside = (np.abs(x - b[0]) < b[3]) & (np.abs(y - b[1]) < b[4]) & (np.abs(z - b[2]) < b[5]) & ~done
        val[inside] = b[6]
        done |= inside
    assert done.all()
    np.testing.assert_array_equal(val, cube[0])


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    cubes = rng.uniform(1, 9, (5, N, N, N)).astype(np.float32)
    G = rng.random((5, 6, N * N)).astype(np.float32)
    L, C = _descriptions(5, 11)
    blk = T.description_blocks(L, C)
    p = str(tmp_path / "poisson.bin")
    T.write_ggft_bin(p, cubes[:3], G[:3], blk[:3])
    T.write_ggft_bin(p, cubes[3:], G[3:], blk[3:], append=True)
    assert os.path.getsize(p) == 8 * (2 + 5 * (N ** 3 + 112 + 6 * N * N))
    c2, b2, g2 = T.read_ggft_bin(p)
    np.testing.assert_array_equal(c2, cubes)
    np.testing.assert_array_equal(b2, blk)
    np.testing.assert_array_equal(g2, G)
    with open(p, "r+b") as f:  # a header of another N is refused
        f.write(struct.pack("<d", 16.0))
    with pytest.raises(ValueError):
        T.read_ggft_bin(p)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference GGFT solver not built (oracle/Makefile)")
@pytest.mark.parametrize("code", [0, 1], ids=["poisson_gft", "gradient_wvtz"])
def test_same_bytes_as_the_reference_solver(tmp_path, code):
    from oracle import oracle as O
    L, C = _descriptions(2, 100 + code)
    cubes, _ = O.voxelize(L, C, N, normalise=False)
    blk = T.description_blocks(L, C)
    ref_file = str(tmp_path / "ref.bin")
    inp = struct.pack("iid", 2, code, 1e-6) + b"".join(
        np.asarray(cubes[i], np.float64).tobytes() + blk[i].tobytes() for i in range(2))
    r = subprocess.run([REF, ref_file, "16"], input=inp, capture_output=True, check=True, timeout=600)
    G = np.frombuffer(r.stdout, np.float64).reshape(2, 6 * N * N + 1)[:, :-1]
    ours = str(tmp_path / "ours.bin")
    T.write_ggft_bin(ours, cubes, G, blk)
    assert open(ours, "rb").read() == open(ref_file, "rb").read()
    c2, b2, g2 = T.read_ggft_bin(ref_file)
    np.testing.assert_array_equal(c2, cubes)
    np.testing.assert_array_equal(b2, blk)
    np.testing.assert_array_equal(g2.reshape(2, -1), G)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_facade_reader(tmp_path):
    from deeprwcap_amd import _lib
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = str(tmp_path / "readers_check")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "readers_check.cpp"), f"-L{libdir}", "-ldnnsolver",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True, timeout=300)
    rng = np.random.default_rng(5)
    cubes = rng.uniform(1, 9, (3, N, N, N))
    G = rng.random((3, 6, N * N))
    L, C = _descriptions(3, 7)
    p = str(tmp_path / "gradient.bin")
    T.write_ggft_bin(p, cubes, G, T.description_blocks(L, C))
    out = str(tmp_path / "read.f32")
    subprocess.run([exe, "--read-bin", p, "3", str(7 * T.GGFT_STRUCTURES), out], check=True, timeout=60)
    v = np.fromfile(out, np.float32)
    np.testing.assert_array_equal(v[:3 * N ** 3], cubes.astype(np.float32).ravel())
    np.testing.assert_array_equal(v[3 * N ** 3:], G.astype(np.float32).ravel())
