"""Python handle of the serving engine (drw_server_* in include/dnnsolver_c.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_int), ("max_inflight", ctypes.c_int), ("hbm_budget_mb", ctypes.c_double),
                ("fdm_tol", ctypes.c_double), ("seed", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("inflight_peak", ctypes.c_uint64),
                ("mean_batch", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("elapsed_s", ctypes.c_double), ("qps", ctypes.c_double)]


class Server:
    """Thread-safe continuous-batching front end for transition queries."""

    def __init__(self, device=0, models_txt=None, max_batch=2048, max_inflight=4, hbm_budget_mb=1024.0,
                 fdm_tol=1e-6, seed=0x5EED5EED):
        cfg = ServerConfig(max_batch, max_inflight, hbm_budget_mb, fdm_tol, seed)
        h = ctypes.c_void_p()
        check(_lib.lib().drw_server_create(device, models_txt.encode() if models_txt else None, ctypes.byref(cfg),
                                           ctypes.byref(h)), "drw_server_create")
        self._h = h

    def submit(self, kind, N, axis, layer, cuboid, qid=0) -> int:
        L = np.ascontiguousarray(layer, np.float32).reshape(-1)
        C = np.ascontiguousarray(cuboid, np.float32).reshape(-1)
        assert L.size == 30 and C.size == 105
        t = ctypes.c_uint64()
        check(_lib.lib().drw_server_submit(self._h, kind, N, axis, L.ctypes.data, C.ctypes.data, qid,
                                           ctypes.byref(t)), "drw_server_submit")
        return t.value

    def wait(self, ticket, kind, timeout_ms=-1):
        out = np.empty(5 if kind == 1 else 3, np.float32)
        check(_lib.lib().drw_server_wait(self._h, ticket, out.ctypes.data, timeout_ms), "drw_server_wait")
        return out

    def stats(self, reset=False) -> dict:
        s = ServerStats()
        check(_lib.lib().drw_server_stats_get(self._h, ctypes.byref(s), int(reset)), "drw_server_stats_get")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().drw_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
