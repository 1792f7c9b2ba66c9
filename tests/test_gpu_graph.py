"""GPU test of the hipGraph-captured query step: every replay (with eager
work and synchronisations in between) equals the eager batch computed with
the same query counters."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def models_txt(tmp_path_factory):
    from deeprwcap_amd import weights as W
    d = np.load(os.path.join(ROOT, "tests", "golden", "nets_golden.npz"))
    states = {n: W.state_from_npz(d, n) for n in W.MODEL_ORDER}
    return W.export_model_dir(str(tmp_path_factory.mktemp("models")), states)


@pytest.mark.parametrize("backend", ["nn", "fdm"])
@pytest.mark.parametrize("kind", [0, 1])
def test_graph_replays_match_eager(backend, kind, models_txt):
    import torch
    from deeprwcap_amd import _lib, structures as S
    from deeprwcap_amd.dnnsolver import DNNSolverGrad, DNNSolverPoisson
    L_ = _lib.lib()
    B = 128 if backend == "nn" else 32
    L, C, _ = S.layered_boxes(B, 23, seed=5 + kind)
    axis = (np.arange(B) % 3).astype(np.int32)
    cls = DNNSolverPoisson if kind == 0 else DNNSolverGrad
    s = cls(0)
    if backend == "nn":
        s.loadModels(models_txt=models_txt)
    else:
        s.use_fdm(23, 1e-8)
    c0 = 1000
    s.set_seed(4321, c0)
    w = 3 if kind == 0 else 5
    dL, dC, dA = torch.from_numpy(L).cuda(), torch.from_numpy(C).cuda(), torch.from_numpy(axis).cuda()
    out = torch.empty((B, w), device="cuda")
    g = ctypes.c_void_p()
    _lib.check(L_.drw_solver_graph_create(s.handle, B, dL.data_ptr(), dC.data_ptr(),
                                          dA.data_ptr() if kind else None, out.data_ptr(), ctypes.byref(g)),
               "graph_create")
    ref_out = torch.empty_like(out)
    try:
        for k in range(4):
            _lib.check(L_.drw_graph_launch(g, None), "graph_launch")
            torch.cuda.synchronize()
            got = out.cpu().numpy().copy()
            x = torch.randn(1 << 16, device="cuda")  # unrelated eager work between replays
            _ = (x * 2).sum().item()
            qid = torch.from_numpy((c0 + k * B + np.arange(B)).astype(np.int64)).cuda()
            _lib.check(L_.drw_solver_sample_ids(s.handle, B, dL.data_ptr(), dC.data_ptr(),
                                                dA.data_ptr() if kind else None, qid.data_ptr(),
                                                ref_out.data_ptr(), None), "sample_ids")
            torch.cuda.synchronize()
            np.testing.assert_array_equal(got, ref_out.cpu().numpy(), err_msg=f"replay {k}")
        cnt = ctypes.c_uint64()
        _lib.check(L_.drw_graph_counter(g, ctypes.byref(cnt)), "graph_counter")
        assert cnt.value == c0 + 4 * B
    finally:
        L_.drw_graph_destroy(g)


def test_graph_replay_on_caller_stream(models_txt):
    """A replay given another stream is ordered around that stream's work: it
    sees the fill queued before it there, the copy queued after it sees its
    result, and eager work on the solver's own stream queued right after it
    (no synchronisation between) neither races with it nor changes it."""
    import torch
    from deeprwcap_amd import _lib, structures as S
    from deeprwcap_amd.dnnsolver import DNNSolverPoisson
    L_ = _lib.lib()
    B = 256
    L, C, _ = S.layered_boxes(B, 23, seed=77)
    s = DNNSolverPoisson(0)
    s.loadModels(models_txt=models_txt)
    c0 = 50000
    s.set_seed(99, c0)
    dL, dC = torch.from_numpy(L).cuda(), torch.from_numpy(C).cuda()
    out = torch.empty((B, 3), device="cuda")
    g = ctypes.c_void_p()
    _lib.check(L_.drw_solver_graph_create(s.handle, B, dL.data_ptr(), dC.data_ptr(), None, out.data_ptr(),
                                          ctypes.byref(g)), "graph_create")
    side = torch.cuda.Stream()
    refs = [torch.empty((B, 3), device="cuda") for _ in range(3)]
    snaps = []
    try:
        for k in range(3):
            with torch.cuda.stream(side):
                out.fill_(float("nan"))
                _lib.check(L_.drw_graph_launch(g, ctypes.c_void_p(side.cuda_stream)), "graph_launch")
                snaps.append(out.clone())
            qid = torch.from_numpy((c0 + k * B + np.arange(B)).astype(np.int64)).cuda()
            torch.cuda.current_stream().synchronize()  # qid is on the device before the solver's stream reads it
            _lib.check(L_.drw_solver_sample_ids(s.handle, B, dL.data_ptr(), dC.data_ptr(), None, qid.data_ptr(),
                                                refs[k].data_ptr(), None), "sample_ids")
        torch.cuda.synchronize()
        for k in range(3):
            got, want = snaps[k].cpu().numpy(), refs[k].cpu().numpy()
            assert np.isfinite(got).all(), f"replay {k} read before it finished"
            np.testing.assert_array_equal(got, want, err_msg=f"replay {k}")
        cnt = ctypes.c_uint64()
        _lib.check(L_.drw_graph_counter(g, ctypes.byref(cnt)), "graph_counter")
        assert cnt.value == c0 + 3 * B
    finally:
        L_.drw_graph_destroy(g)
