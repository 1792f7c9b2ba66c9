"""Pins of the FP64 FDM oracle (CPU only).

Each check is chosen so that a plausible slip in the restatement (a wrong
face weight, a dropped boundary term, a transposed face layout, a wrong
even-N source) fails at least one of them:

* GGFT golden vectors from the reference solver compiled from its sources
  (N = 16, 21 and 23, GFT / WVTZ / WVTX / WVTY, tolerance 1e-13; N = 16
  takes the reference's even-N source and gradient stencils);
* homogeneous eps: separable DST-II closed form, odd and even N;
* z-layered eps: DST-II in x, y and a tridiagonal solve per mode in z;
* tiny N: a dense solve of GGFT's own assembled matrix A = L D^-1;
* invariants: sum G = 1, G >= 0 (Poisson), sum G^z = 0 (gradient), scale
  invariance and the z-reflection symmetry.
"""
import numpy as np
import pytest


def _dst_basis(N):
    j = np.arange(N) + 0.5
    m = np.arange(1, N + 1)
    V = np.sin(np.pi * np.outer(j, m) / N)
    V /= np.linalg.norm(V, axis=0)
    lam = 2.0 - 2.0 * np.cos(np.pi * m / N)
    return V, lam


def _homog_closed_form(N, e):
    """y = L^-1 e for eps = 1 by the DST-II eigen-decomposition of
    tridiag(-1, 2, -1) with 3 on both end diagonals (Dirichlet at half a cell)."""
    V, lam = _dst_basis(N)
    E = e.reshape(N, N, N)  # (z, y, x)
    Eh = np.einsum("zyx,xa,yb,zc->cba", E, V, V, V)
    Lam = lam[:, None, None] + lam[None, :, None] + lam[None, None, :]
    Y = np.einsum("cba,xa,yb,zc->zyx", Eh / Lam, V, V, V)
    return Y.reshape(-1)


def _layered_closed_form(N, epsz, e):
    """z-layered eps: DST-II in x and y, then per (a, b) mode a tridiagonal
    solve in z with harmonic-mean couplings and 2 eps boundary terms."""
    V, lam = _dst_basis(N)
    E = e.reshape(N, N, N)
    Eh = np.einsum("zyx,xa,yb->zba", E, V, V)
    w = 2 * epsz[:-1] * epsz[1:] / (epsz[:-1] + epsz[1:])
    Lz = np.zeros((N, N))
    for k in range(N - 1):
        Lz[k, k] += w[k]
        Lz[k + 1, k + 1] += w[k]
        Lz[k, k + 1] -= w[k]
        Lz[k + 1, k] -= w[k]
    Lz[0, 0] += 2 * epsz[0]
    Lz[-1, -1] += 2 * epsz[-1]
    Yh = np.empty_like(Eh)
    for a in range(N):
        for b in range(N):
            M = Lz + np.diag((lam[a] + lam[b]) * epsz)
            Yh[:, b, a] = np.linalg.solve(M, Eh[:, b, a])
    Y = np.einsum("zba,xa,yb->zyx", Yh, V, V)
    return Y.reshape(-1)


def _G_from_y(orc, eps, y):
    N = eps.shape[-1]
    cells = orc.panel_cells(N)
    return 2.0 * eps.reshape(-1)[cells] * y[cells]


def _relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_golden_ggft(orc, golden_ggft_n):
    g = golden_ggft_n
    N = int(g["N"])
    cubes, _ = orc.voxelize(g["layers"], g["cuboids"], N, normalise=False)
    for b in range(len(g["codes"])):
        G, it, rr = orc.fdm_solve(cubes[b].astype(np.float64), int(g["codes"][b]), tol=1e-14)
        assert it > 0
        assert _relerr(G, g["G"][b]) < 1e-10, (b, _relerr(G, g["G"][b]))


@pytest.mark.parametrize("N", [5, 6, 9, 16])
@pytest.mark.parametrize("code", [0, 1, 2, 3])
def test_homogeneous_closed_form(orc, N, code):
    eps = np.ones((N, N, N))
    e = orc.rhs(eps, code)
    y = _homog_closed_form(N, e)
    G_cf = _G_from_y(orc, eps, y)
    G, it, _ = orc.fdm_solve(eps, code, tol=1e-14)
    assert it > 0
    assert _relerr(G, G_cf) < 1e-11


@pytest.mark.parametrize("N", [7, 8])
@pytest.mark.parametrize("code", [0, 1])
def test_layered_closed_form(orc, N, code):
    rng = np.random.default_rng(N + 10 * code)
    epsz = rng.uniform(1, 8, N)
    eps = np.broadcast_to(epsz[:, None, None], (N, N, N)).copy()
    e = orc.rhs(eps, code)
    y = _layered_closed_form(N, epsz, e)
    np.testing.assert_allclose(orc.apply_L(eps, y), e, atol=1e-12)
    G, it, _ = orc.fdm_solve(eps, code, tol=1e-14)
    assert _relerr(G, _G_from_y(orc, eps, y)) < 1e-11


@pytest.mark.parametrize("N", [3, 4, 5])
@pytest.mark.parametrize("code", [0, 1, 2, 3])
def test_dense_ggft_matrix(orc, N, code):
    """GGFT solves A x = e with its own A; G = x on the boundary cells
    (ggft_solver.cpp:552-596).  The oracle solves L y = e with x = 2 eps y."""
    rng = np.random.default_rng(100 * N + code)
    eps = rng.uniform(1, 8, (N, N, N))
    A = orc.ggft_dense_A(eps)
    x = np.linalg.solve(A, orc.rhs(eps, code))
    G_dense = x[orc.panel_cells(N)]
    G, it, _ = orc.fdm_solve(eps, code, tol=1e-14)
    assert _relerr(G, G_dense) < 1e-11


@pytest.mark.parametrize("N", [8, 9])
def test_invariants(orc, N):
    rng = np.random.default_rng(N)
    eps = rng.uniform(1, 8, (N, N, N))
    G, _, _ = orc.fdm_solve(eps, 0, tol=1e-14)
    assert abs(G.sum() - 1) < 1e-12 and G.min() >= 0
    for code in (1, 2, 3):
        Gg, _, _ = orc.fdm_solve(eps, code, tol=1e-14)
        assert abs(Gg.sum()) < 1e-11 * np.abs(Gg).sum()
    # scale invariance
    G2, _, _ = orc.fdm_solve(3.7 * eps, 0, tol=1e-14)
    assert _relerr(G2, G) < 1e-11
    # z reflection: faces 0 and 1 swap; for odd N the z-gradient kernel
    # changes sign.  (The even-N WVTZ source of ggft_solver.cpp:203-225 is
    # antisymmetric only when both 4-cell groups have equal eps sums.)
    epsr = eps[::-1].copy()
    NN = N * N
    Gr, _, _ = orc.fdm_solve(epsr, 0, tol=1e-14)
    assert _relerr(Gr[:NN], G[NN:2 * NN]) < 1e-11
    if N % 2 == 0:
        return
    Gz, _, _ = orc.fdm_solve(eps, 1, tol=1e-14)
    Gzr, _, _ = orc.fdm_solve(epsr, 1, tol=1e-14)
    assert _relerr(-Gzr[:NN], Gz[NN:2 * NN]) < 1e-10


def test_even_n_sources_sum(orc):
    rng = np.random.default_rng(5)
    eps = rng.uniform(1, 8, (16, 16, 16))
    assert abs(orc.rhs(eps, 0).sum() - 1) < 1e-14
    for code in (1, 2, 3):
        e = orc.rhs(eps, code)
        assert np.count_nonzero(e) == 8 and abs(e.sum()) < 1e-12 * np.abs(e).sum()
