"""Full-size mixed walker batch: 1,048,576 transition queries, N uniform over
{16, 21, 23}, half Poisson and half gradient (axis uniform), one Philox
counter each; N = 23 runs on the networks, N = 16 / 21 on the FDM path
(batched PCG + flat inverse-CDF over the 6 N^2 tiles).

Size-independent properties, on every query:
* the hop is a panel centre of the cube surface: exactly one coordinate is
  +-1, the others are (2i + 1)/N - 1 in float32 (locate_index_simple_kernel,
  inference_cpp/src/cuda_kernels.cu:219-264);
* gradient queries: W sign finite and non-zero, centre permittivity finite
  and positive;
* batch independence: a random subset re-run in batches of its own gives
  bit-identical results;
and, on a subset of the FDM queries, hop draws bit-exact against the CPU
sampler given the GPU's FDM kernels and the query's uniforms.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N_QUERIES = 1 << 20
SEED = 0x5EED


@pytest.fixture(scope="module")
def models_txt(tmp_path_factory):
    from deeprwcap_amd import weights as W
    d = np.load(os.path.join(ROOT, "tests", "golden", "nets_golden.npz"))
    states = {n: W.state_from_npz(d, n) for n in W.MODEL_ORDER}
    return W.export_model_dir(str(tmp_path_factory.mktemp("models")), states)


@pytest.fixture(scope="module")
def full_run(models_txt):
    from deeprwcap_amd.mixed import MixedRunner, make_queries
    q = make_queries(N_QUERIES, seed=3)
    runner = MixedRunner(models_txt, seed=SEED)
    out, groups, total = runner.run(q)
    print(f"{N_QUERIES} mixed queries in {total:.3f} s ({N_QUERIES / total:,.0f} queries/s): {groups}", flush=True)
    return q, runner, out


def _panel_centres_ok(p, N):
    face = np.abs(p) == 1.0
    if not (face.sum(1) == 1).all():
        return False
    allowed = {}
    for n in np.unique(N):
        allowed[int(n)] = np.sort((np.float32(2 * np.arange(n) + 1) / np.float32(n) - np.float32(1)).astype(np.float32))
    ok = np.ones(p.shape[0], bool)
    for n, vals in allowed.items():
        rows = N == n
        other = p[rows][~face[rows]].reshape(-1, 2)
        ok[rows] = np.isin(other, vals).all(1)
    return ok.all()


def test_full_batch_properties(full_run):
    q, _, out = full_run
    assert np.isfinite(out[:, :3]).all()
    assert _panel_centres_ok(out[:, :3], q["N"])
    g = q["kind"] == 1
    assert np.isfinite(out[g, 3:]).all()
    assert (out[g, 3] != 0).all()
    assert (out[g, 4] > 0).all()


def test_batch_independence(full_run):
    q, runner, out = full_run
    rng = np.random.default_rng(11)
    sub = np.sort(rng.choice(N_QUERIES, 3000, replace=False))
    again, _, _ = runner.run(q, subset=sub)
    np.testing.assert_array_equal(again[sub], out[sub])


def test_fdm_draws_bitexact(full_run, orc):
    import torch
    from deeprwcap_amd import device as D
    q, _, out = full_run
    rng = np.random.default_rng(12)
    for kind in (0, 1):
        idx = rng.choice(np.nonzero((q["N"] != 23) & (q["kind"] == kind))[0], 24, replace=False)
        Ns = q["N"][idx]
        pk = D.PackedCubes(Ns)
        cubes = [orc.voxelize(q["layers"][i:i + 1], q["cuboids"][i:i + 1], int(q["N"][i]), normalise=False)[0][0]
                 for i in idx]
        pk.set_cubes(cubes)
        codes = [0 if kind == 0 else {0: 2, 1: 3, 2: 1}[int(q["axis"][i])] for i in idx]
        D.fdm_solve(pk, codes, tol=1e-6, want_stats=False)
        G = pk.G.cpu().numpy()
        torch.cuda.synchronize()
        for j, i in enumerate(idx):
            n = int(Ns[j])
            g = G[pk.G_off[j]:pk.G_off[j] + 6 * n * n]
            u = orc.query_uniforms(SEED, int(q["qid"][i]))
            t = orc.cdf_draw(g, int(u[0]))
            f, tt = divmod(t, n * n)
            np.testing.assert_array_equal(out[i, :3], orc.locate(n, tt, f), err_msg=f"query {i} N {n} kind {kind}")
            if kind == 1:
                assert out[i, 3] == pytest.approx(np.abs(g).sum() * np.sign(g[t]), rel=1e-6)
                assert out[i, 4] == pytest.approx(orc.center_eps(cubes[j].astype(np.float64), codes[j]), rel=1e-6)
