"""GPU tests of failure handling on the query path (the acceptance and
fallback of ggft_solver::solve, ggft/src/ggft_solver.cpp:565-626), each
failure forced from the host so that every kernel runs to its normal end:

* FDM backend: the first solve is capped (drw_solver_set_fdm_limits: one
  refinement round of three FP32 iterations), so no cube converges; every
  cube is then solved again in FP64 mode, the counters say so, and the hops
  equal those of an uncapped solver at tol 1e-12 on the same counters.
* Network backend: models whose selector head bias is NaN make every face
  draw fail; every query goes to the FDM fallback, and the results equal the
  FDM backend's on the same counters bit for bit (same voxel values, same
  solve, same draw).
* Normal batches trigger neither path.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "models", "models.txt")
SEED, C0 = 0xFA11BAC, 777


def _failures(s):
    from deeprwcap_amd import _lib
    v = [ctypes.c_uint64() for _ in range(3)]
    _lib.check(_lib.lib().drw_solver_failures(s.handle, *[ctypes.byref(x) for x in v]), "failures")
    return tuple(x.value for x in v)


def _batch(B, seed):
    from deeprwcap_amd import structures as S
    L, C, _ = S.layered_boxes(B, 23, seed=seed)
    axis = (np.arange(B) % 3).astype(np.int32)
    return L, C, axis


def _run(s, L, C, axis, kind):
    s.set_seed(SEED, C0)
    out = s.sample(L.reshape(-1), C.reshape(-1), axis if kind else None)
    return np.asarray(out, np.float32).reshape(len(L), 5 if kind else 3)


@pytest.mark.parametrize("kind", [0, 1], ids=["poisson", "gradient"])
def test_fdm_capped_solve_is_resolved(kind):
    import torch  # noqa: F401
    from deeprwcap_amd import _lib
    from deeprwcap_amd.dnnsolver import DNNSolverGrad, DNNSolverPoisson
    cls = DNNSolverGrad if kind else DNNSolverPoisson
    B = 64
    L, C, axis = _batch(B, 501 + kind)
    capped, ref = cls(0), cls(0)
    capped.use_fdm(23, 1e-6)
    ref.use_fdm(23, 1e-12)
    _lib.check(_lib.lib().drw_solver_set_fdm_limits(capped.handle, 1, 3), "set_fdm_limits")
    print(f"capped FDM kind {kind} B {B}", flush=True)
    got = _run(capped, L, C, axis, kind)
    want = _run(ref, L, C, axis, kind)
    resolved, failed, nn = _failures(capped)
    assert (resolved, failed, nn) == (B, 0, 0), (resolved, failed, nn)
    assert _failures(ref) == (0, 0, 0)
    assert np.isfinite(got).all()
    np.testing.assert_array_equal(got[:, :3], want[:, :3])
    if kind:
        np.testing.assert_allclose(got[:, 3], want[:, 3], rtol=1e-9)
        np.testing.assert_array_equal(got[:, 4], want[:, 4])


@pytest.mark.parametrize("kind", [0, 1], ids=["poisson", "gradient"])
def test_network_failures_go_to_fdm(kind, tmp_path):
    import torch  # noqa: F401
    from deeprwcap_amd import weights as W
    from deeprwcap_amd.dnnsolver import DNNSolverGrad, DNNSolverPoisson
    states = {n: W.read_drwnet(os.path.join(ROOT, "models", f"{n}.drwnet")) for n in W.MODEL_ORDER}
    sel = "GradientSelectorWeight" if kind else "PoissonSelector"
    last = max(k for k in states[sel] if k.startswith("head.") and k.endswith(".bias"))
    states[sel][last] = np.full_like(states[sel][last], np.nan)
    broken = W.export_model_dir(str(tmp_path / "broken"), states)
    cls = DNNSolverGrad if kind else DNNSolverPoisson
    B = 96
    L, C, axis = _batch(B, 601 + kind)
    nn, fdm, good = cls(0), cls(0), cls(0)
    nn.loadModels(models_txt=broken)
    fdm.use_fdm(23, 1e-6)
    good.loadModels(models_txt=MODELS)
    print(f"network fallback kind {kind} B {B} (NaN {sel} {last})", flush=True)
    got = _run(nn, L, C, axis, kind)
    want = _run(fdm, L, C, axis, kind)
    assert _failures(nn) == (0, 0, B)
    assert np.isfinite(got).all()
    np.testing.assert_array_equal(got, want)
    # a normal batch takes neither path
    ok = _run(good, L, C, axis, kind)
    assert np.isfinite(ok).all()
    assert _failures(good) == (0, 0, 0)
