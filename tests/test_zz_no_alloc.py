"""Allocation-free network query through the C ABI.

drw_solver_sample_device_noalloc must give the hops of
drw_solver_sample_device bit for bit (same kernels, same Philox counters),
on two calls in a row of which the second is smaller, for both query kinds,
advance the solver's counter the same way, and refuse batches beyond its
maximum and the FDM backend with an error code.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "models", "models.txt").encode()


@pytest.mark.parametrize("kind", [0, 1], ids=["poisson", "gradient"])
def test_noalloc_matches_sample_device(kind):
    import torch
    from deeprwcap_amd import _lib, structures
    L = _lib.lib()
    assert torch.cuda.is_available()
    s = ctypes.c_void_p()
    _lib.check(L.drw_solver_create(kind, 0, ctypes.byref(s)), "solver_create")
    try:
        _lib.check(L.drw_solver_load_models(s, MODELS), "load_models")
        width = 5 if kind else 3
        for call, B in enumerate((4096, 1000)):
            print(f"noalloc kind {kind} call {call} B {B}", flush=True)
            Ls, Cs, _ = structures.layered_boxes(B, 23, seed=50 + call)
            dL = torch.from_numpy(Ls).cuda()
            dC = torch.from_numpy(Cs).cuda()
            ax = torch.randint(0, 3, (B,), dtype=torch.int32, device="cuda") if kind else None
            ref = torch.empty((B, width), device="cuda")
            got = torch.full((B, width), float("nan"), device="cuda")
            c0 = 100000 * (call + 1) + 7
            _lib.check(L.drw_solver_set_seed(s, 99, c0), "set_seed")
            _lib.check(L.drw_solver_sample_device(s, B, dL.data_ptr(), dC.data_ptr(), _lib.ptr(ax), ref.data_ptr(),
                                                  None), "sample_device")
            _lib.check(L.drw_solver_set_seed(s, 99, c0), "set_seed")
            _lib.check(L.drw_solver_sample_device_noalloc(s, B, dL.data_ptr(), dC.data_ptr(), _lib.ptr(ax),
                                                          got.data_ptr(), None), "sample_device_noalloc")
            torch.cuda.synchronize()
            ctr = ctypes.c_uint64()
            _lib.check(L.drw_solver_get_counter(s, ctypes.byref(ctr)), "get_counter")
            assert ctr.value == c0 + B
            r, g = ref.cpu().numpy(), got.cpu().numpy()
            assert np.isfinite(r).all()
            assert np.array_equal(r, g), f"call {call}: {int((r != g).any(1).sum())} of {B} queries differ"
        # beyond the maximum: a clear error, nothing launched
        rc = L.drw_solver_sample_device_noalloc(s, 8193, dL.data_ptr(), dC.data_ptr(), _lib.ptr(ax),
                                                got.data_ptr(), None)
        assert rc == -2 and b"exceeds" in L.drw_last_error()
        # the FDM backend allocates: refused
        _lib.check(L.drw_solver_use_fdm(s, 23, 1e-6), "use_fdm")
        rc = L.drw_solver_sample_device_noalloc(s, 16, dL.data_ptr(), dC.data_ptr(), _lib.ptr(ax),
                                                got.data_ptr(), None)
        assert rc == -4 and b"network backend" in L.drw_last_error()
    finally:
        L.drw_solver_destroy(s)
