"""Pins of the oracle's voxelizer, index remaps, locate table and sampler.

* axis pre-rotation against torch.rot90 (library routine);
* face-rotation and post-process tables against FDM equivariance: solving
  the rotated cube must reproduce the kernel of the selected face;
* locate against the GGFT panel numbering (panel centre of the boundary cell);
* Philox4x32-10 against the published known-answer vectors (Random123);
* the fixed-point inverse-CDF draw against its distribution (chi^2), edge
  uniforms and zero-weight bins;
* voxelization against structural properties (layer counts, box volumes,
  normalisation, first-match priority, padding).
"""
import numpy as np
import pytest


def _relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_rotate_matches_torch_rot90(orc, axis):
    import torch
    N = 7
    cube = np.random.default_rng(axis).standard_normal((N, N, N)).astype(np.float32)
    t = torch.from_numpy(cube)[None, None, None]  # [B,1,1,z,y,x]
    if axis == 0:
        ref = t.rot90(-1, dims=(-3, -1))
    elif axis == 1:
        ref = t.rot90(-1, dims=(-3, -2))
    else:
        ref = t
    np.testing.assert_array_equal(orc.axis_rotate(cube, axis), ref[0, 0, 0].numpy())


@pytest.mark.parametrize("N", [7, 8])
def test_poisson_rotation_equivariance(orc, N):
    """FDM(rotate_poisson_f(eps)) face 0 == FDM(eps) face f (t = a + N b)."""
    rng = np.random.default_rng(N)
    eps = rng.uniform(1, 8, (N, N, N)).astype(np.float32)
    G, _, _ = orc.fdm_solve(eps.astype(np.float64), 0, tol=1e-13)
    NN = N * N
    for f in range(6):
        rot = orc.rotate_poisson(eps, f)
        assert sorted(rot.ravel()) == sorted(eps.ravel())
        Gr, _, _ = orc.fdm_solve(rot.astype(np.float64), 0, tol=1e-13)
        assert _relerr(Gr[:NN], G[f * NN:(f + 1) * NN]) < 1e-9, f


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_gradient_tables_equivariance(orc, axis):
    """The gradient query's chain, with FDM kernels in place of the networks:
    pre-rotate by axis; for face f take the target face of the face-rotated
    cube (face 1 for f in {0,1} as Gradient1Predictor, face 2 for f >= 2 as
    Gradient2Predictor); post-process by (axis, f) -> equals the original
    cube's axis-gradient kernel on face remap(f, axis)."""
    N = 7
    NN = N * N
    code = {0: 2, 1: 3, 2: 1}[axis]
    rng = np.random.default_rng(30 + axis)
    eps = rng.uniform(1, 8, (N, N, N)).astype(np.float32)
    Ga, _, _ = orc.fdm_solve(eps.astype(np.float64), code, tol=1e-13)
    cube_r = orc.axis_rotate(eps, axis)
    remap = {0: {0: 4, 1: 5, 4: 1, 5: 0}, 1: {0: 2, 1: 3, 2: 1, 3: 0}, 2: {}}[axis]
    for f in range(6):
        rot = orc.rotate_gradient(cube_r, f)
        Gz, _, _ = orc.fdm_solve(rot.astype(np.float64), 1, tol=1e-13)
        tgt = 1 if f < 2 else 2
        pred = Gz[tgt * NN:(tgt + 1) * NN].astype(np.float32).reshape(N, N)
        post = orc.post_process(pred, f, axis).reshape(-1)
        fo = remap.get(f, f)
        # locate: the tile's point must be the panel centre on original face fo
        assert _relerr(post, Ga[fo * NN:(fo + 1) * NN]) < 1e-6, (axis, f)
        for t in (0, 5, NN - 1):
            p = orc.locate(N, t, f, axis)
            q = orc.locate(N, t, fo, -1)
            np.testing.assert_array_equal(p, q)


@pytest.mark.parametrize("N", [7, 8, 23])
def test_locate_matches_ggft_panels(orc, N):
    cells = orc.panel_cells(N)
    NN = N * N
    for f in range(6):
        for t in range(0, NN, 5):
            c = cells[f * NN + t]
            i, j, k = c % N, (c // N) % N, c // NN
            ctr = np.array([(2 * i + 1) / N - 1, (2 * j + 1) / N - 1, (2 * k + 1) / N - 1])
            ax, sgn = [(2, -1), (2, 1), (1, -1), (1, 1), (0, -1), (0, 1)][f]
            ctr[ax] = sgn
            np.testing.assert_allclose(orc.locate(N, t, f), ctr, atol=1e-6)


def test_philox_known_answers(orc):
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, out in kat:
        assert list(orc.philox(ctr, key)) == out


def test_cdf_draw_edges(orc):
    w = np.array([0, 0, 3.0, 0, 1.0, 0], np.float32)
    assert orc.cdf_draw(w, 0) == 2
    assert orc.cdf_draw(w, 0xFFFFFFFF) == 4
    assert orc.cdf_draw(-w, 0) == 2  # magnitudes
    assert orc.cdf_draw(np.zeros(5, np.float32), 7) == -1
    one = np.zeros(529, np.float32)
    one[300] = 1e-30
    assert all(orc.cdf_draw(one, u) == 300 for u in (0, 1 << 31, 0xFFFFFFFF))
    assert orc.cdf_draw(w.astype(np.float64), 0xFFFFFFFF) == 4


def test_cdf_draw_distribution(orc):
    rng = np.random.default_rng(3)
    w = rng.random(40).astype(np.float32) ** 3
    w[[3, 17]] = 0
    n = 40000
    us = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    counts = np.bincount([orc.cdf_draw(w, int(u)) for u in us], minlength=40)
    assert counts[3] == 0 and counts[17] == 0
    p = w / w.sum()
    m = p > 0
    chi2 = ((counts[m] - n * p[m]) ** 2 / (n * p[m])).sum()
    assert chi2 < 38 + 5 * np.sqrt(2 * 38)


def test_voxelize_properties(orc):
    from deeprwcap_amd import structures as S
    N = 23
    L, C = S.empty(3)
    # case 0: two z-layers, interface at grid face 9 -> 9 planes eps 2, 14 planes eps 8
    L[0, 0] = (-1.5, 2.0)
    L[0, 1] = (-1 + 2 * 9 / N, 8.0)
    # case 1: background 1, box on cells 2..5 x 3..9 x 0..22 with eps 4, then a
    # second box (lower priority) overlapping it
    L[1, 0] = (-1.5, 1.0)
    C[1, 0] = (-1 + 4 / N, -1 + 6 / N, -1.5, -1 + 12 / N, -1 + 20 / N, 1.5, 4.0)
    C[1, 1] = (-1.5, -1.5, -1.5, 0.0, 0.0, 0.0, 3.0)
    # case 2: padding in the middle of the layer list ends it
    L[2, 0] = (-1.5, 5.0)
    L[2, 1, 0] = S.PAD
    L[2, 2] = (0.0, 9.0)
    cube, mx = orc.voxelize(L, C, N, normalise=True)
    assert mx[0] == 8.0 and mx[1] == 4.0 and mx[2] == 9.0  # max includes unused entries
    assert (cube[0, :9] == np.float32(2.0) / 8).all() and (cube[0, 9:] == 1.0).all()
    assert (cube[1][:, 3:10, 2:6] == 1.0).sum() == 4 * 7 * N
    assert (cube[1] == np.float32(3.0) / 4).sum() == 12 * 12 * 12 - 4 * 7 * 12
    assert (cube[2] == np.float32(5.0) / 9).all()
    raw, _ = orc.voxelize(L, C, N, normalise=False)
    assert (raw[1][:, 3:10, 2:6] == 4.0).all()
