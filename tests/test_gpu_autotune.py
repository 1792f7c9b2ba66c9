"""The autotuner's launch-shape search on the GPU, and agreement of the face
kernels' two launch shapes (256 and 512 threads per query and face)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "models", "models.txt")


def _set_all(lib, threads):
    from deeprwcap_amd import _lib
    for w in range(4):
        _lib.check(lib.drw_set_launch_shape(w, -1, threads), "set_launch_shape")


@pytest.mark.parametrize("kind", [0, 1])
def test_shapes_agree(kind):
    """Both shapes compute the same network: kernels, selector outputs and
    hops bit-identical (the output normalisation's sum runs in the same order
    at both widths)."""
    import torch
    from deeprwcap_amd import _lib, structures as S
    from deeprwcap_amd.dnnsolver import DNNSolverGrad, DNNSolverPoisson
    lib = _lib.lib()
    B = 256
    L, C, _ = S.layered_boxes(B, 23, seed=300 + kind)
    axis = (np.arange(B) % 3).astype(np.int32)
    dL, dC, dA = torch.from_numpy(L).cuda(), torch.from_numpy(C).cuda(), torch.from_numpy(axis).cuda()
    s = (DNNSolverPoisson if kind == 0 else DNNSolverGrad)(0)
    s.loadModels(models_txt=MODELS)
    res = {}
    try:
        for th in (256, 512):
            _set_all(lib, th)
            probs = torch.empty((B, 6), device="cuda")
            wts = torch.empty(B, device="cuda")
            kern = torch.empty((B, 6, 529), device="cuda")
            _lib.check(lib.drw_solver_eval_kernels(s.handle, B, dL.data_ptr(), dC.data_ptr(),
                                                   dA.data_ptr() if kind else None, probs.data_ptr(),
                                                   wts.data_ptr(), kern.data_ptr(), None), "eval_kernels")
            s.set_seed(11, 5000)
            out = torch.empty((B, 3 if kind == 0 else 5), device="cuda")
            _lib.check(lib.drw_solver_sample_device(s.handle, B, dL.data_ptr(), dC.data_ptr(),
                                                    dA.data_ptr() if kind else None, out.data_ptr(), None),
                       "sample_device")
            torch.cuda.synchronize()
            res[th] = [t.cpu().numpy() for t in (probs, wts, kern, out)]
    finally:
        _set_all(lib, 0)
    (p0, w0, k0, o0), (p1, w1, k1, o1) = res[256], res[512]
    np.testing.assert_array_equal(p0, p1)
    if kind == 1:  # Poisson queries have no weight output
        np.testing.assert_array_equal(w0, w1)
    np.testing.assert_array_equal(k0, k1)
    np.testing.assert_array_equal(o0, o1)


def test_tune_shapes():
    """The search times both shapes of every face kernel and installs rules
    that begin at 0 workgroups."""
    from deeprwcap_amd import _lib, autotune
    lib = _lib.lib()
    try:
        res = autotune.tune_shapes(MODELS, ladder=(64, 4096), log=lambda m: print(m, flush=True))
        rules = res["launch_shapes"]
        assert set(rules) == {"face_poisson", "face_grad1", "face_grad2", "poisson_fused"}
        for name, r in rules.items():
            shapes = (1, 256, 512) if name == "poisson_fused" else (256, 512)
            assert r[0][0] == 0 and all(th in shapes for _, th in r)
            assert [m for m, _ in r] == sorted(m for m, _ in r)
            which = autotune.SHAPE_KERNELS[name]
            for B in (64, 4096):
                want = [th for m, th in r if m <= B][-1]
                assert lib.drw_get_launch_shape(which, B) == want
            times = res["shape_ms"][name]
            want_keys = {str(x) for x in shapes}
            assert set(times) == {"64", "4096"} and all(set(t) == want_keys for t in times.values())
            assert all(0 < v < 50 for t in times.values() for v in t.values())
    finally:
        _set_all(lib, 0)


def test_fused_poisson_query_matches_separate():
    """The fused Poisson query kernel (256 and 512 threads) gives exactly the
    separate kernels' results: hops, drawn faces, face probabilities and the
    drawn face's kernel."""
    import torch
    from deeprwcap_amd import _lib, structures as S
    from deeprwcap_amd.dnnsolver import DNNSolverPoisson
    lib = _lib.lib()
    B = 1536
    L, C, _ = S.layered_boxes(B, 23, seed=310)
    dL, dC = torch.from_numpy(L).cuda(), torch.from_numpy(C).cuda()
    s = DNNSolverPoisson(0)
    s.loadModels(models_txt=MODELS)
    res = {}
    try:
        for th in (1, 256, 512):
            _lib.check(lib.drw_set_launch_shape(3, -1, th), "set_launch_shape")
            s.set_seed(77, 9000)
            out = torch.empty((B, 3), device="cuda")
            face = torch.empty(B, dtype=torch.int32, device="cuda")
            probs = torch.empty((B, 6), device="cuda")
            kern = torch.empty((B, 529), device="cuda")
            _lib.check(lib.drw_solver_sample_ex(s.handle, B, dL.data_ptr(), dC.data_ptr(), None, out.data_ptr(),
                                                face.data_ptr(), probs.data_ptr(), None, kern.data_ptr(), None),
                       "sample_ex")
            torch.cuda.synchronize()
            res[th] = [t.cpu().numpy() for t in (out, face, probs, kern)]
    finally:
        _set_all(lib, 0)
    assert np.isfinite(res[1][0]).all()
    for th in (256, 512):
        for name, a, b in zip(("out", "face", "probs", "kern"), res[1], res[th]):
            np.testing.assert_array_equal(a, b, err_msg=f"fused {th} threads: {name}")
