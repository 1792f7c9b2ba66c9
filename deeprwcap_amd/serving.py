"""Python handle of the serving engine (drw_server_* in include/dnnsolver_c.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_int), ("max_inflight", ctypes.c_int), ("hbm_budget_mb", ctypes.c_double),
                ("fdm_tol", ctypes.c_double), ("seed", ctypes.c_uint64), ("resident", ctypes.c_int),
                ("resident_threads", ctypes.c_int), ("resident_idle_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("inflight_peak", ctypes.c_uint64),
                ("mean_batch", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("elapsed_s", ctypes.c_double), ("qps", ctypes.c_double)]


class LoadConfig(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int), ("max_outstanding", ctypes.c_int), ("rate_qps", ctypes.c_double),
                ("duration_s", ctypes.c_double), ("seed", ctypes.c_uint64)]


class Geometry(ctypes.Structure):
    _fields_ = [("window", ctypes.c_double * 6), ("n_layers", ctypes.c_int), ("layer_z", ctypes.c_void_p),
                ("layer_eps", ctypes.c_void_p), ("n_boxes", ctypes.c_int), ("cond_box", ctypes.c_void_p),
                ("cond_id", ctypes.c_void_p), ("n_media", ctypes.c_int), ("media_box", ctypes.c_void_p),
                ("media_eps", ctypes.c_void_p)]


class WalkConfig(ctypes.Structure):
    _fields_ = [("walks", ctypes.c_int), ("threads", ctypes.c_int), ("concurrent", ctypes.c_int), ("N", ctypes.c_int),
                ("stop_dist", ctypes.c_double), ("gauss_frac", ctypes.c_double), ("max_cube", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class Server:
    """Thread-safe continuous-batching front end for transition queries."""

    def __init__(self, device=0, models_txt=None, max_batch=2048, max_inflight=4, hbm_budget_mb=1024.0,
                 fdm_tol=1e-6, seed=0x5EED5EED, resident=False, resident_threads=512, resident_idle_ms=20.0):
        """resident: N = 23 queries go to the resident loop (a persistent
        kernel reading a pinned request ring) instead of batches."""
        cfg = ServerConfig(max_batch, max_inflight, hbm_budget_mb, fdm_tol, seed, int(bool(resident)),
                           resident_threads, resident_idle_ms)
        h = ctypes.c_void_p()
        check(_lib.lib().drw_server_create(device, models_txt.encode() if models_txt else None, ctypes.byref(cfg),
                                           ctypes.byref(h)), "drw_server_create")
        self._h = h

    @classmethod
    def from_tuning(cls, path, device=0, models_txt=None, **kw):
        """Server with the batch shape chosen by deeprwcap_amd.autotune (whose
        launch shapes are applied to the whole process)."""
        import json
        from . import autotune
        with open(path) as f:
            t = json.load(f)
        autotune.apply(t)
        kw.setdefault("max_batch", int(t["max_batch"]))
        kw.setdefault("max_inflight", int(t.get("max_inflight", 4)))
        return cls(device, models_txt, **kw)

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

    def submit_many(self, kinds, Ns, axes, layers, cuboids, qids=None) -> np.ndarray:
        """Submit n queries in one call; returns their tickets (uint64 [n])."""
        K = np.ascontiguousarray(kinds, np.int32)
        n = K.size
        Nn = np.ascontiguousarray(Ns, np.int32)
        A = np.ascontiguousarray(axes, np.int32)
        L = np.ascontiguousarray(layers, np.float32).reshape(n, 30)
        C = np.ascontiguousarray(cuboids, np.float32).reshape(n, 105)
        assert Nn.size == n and A.size == n
        Q = None if qids is None else np.ascontiguousarray(qids, np.uint64)
        t = np.zeros(n, np.uint64)
        check(_lib.lib().drw_server_submit_many(self._h, n, K.ctypes.data, Nn.ctypes.data, A.ctypes.data,
                                                L.ctypes.data, C.ctypes.data, None if Q is None else Q.ctypes.data,
                                                t.ctypes.data), "drw_server_submit_many")
        return t

    def wait_many(self, tickets, timeout_ms=-1) -> np.ndarray:
        """Wait for tickets in order; returns float32 [n, 5] (Poisson rows use
        the first 3 columns)."""
        T = np.ascontiguousarray(tickets, np.uint64)
        out = np.zeros((T.size, 5), np.float32)
        check(_lib.lib().drw_server_wait_many(self._h, T.size, T.ctypes.data, out.ctypes.data, timeout_ms),
              "drw_server_wait_many")
        return out

    def stats(self, reset=False) -> dict:
        s = ServerStats()
        check(_lib.lib().drw_server_stats_get(self._h, ctypes.byref(s), int(reset)), "drw_server_stats_get")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def resident_stats(self) -> dict:
        """Resident-loop counters (drw_server_resident_stats)."""
        out = (ctypes.c_uint64 * 10)()
        check(_lib.lib().drw_server_resident_stats(self._h, out), "drw_server_resident_stats")
        keys = ("launches", "served_poisson", "served_gradient", "scouts", "exits_idle", "exits_stop", "claimed",
                "published", "grid", "wall_clock_khz")
        return dict(zip(keys, (int(v) for v in out)))

    def run_load(self, layers, cuboids, kinds, Ns, axes, threads=64, rate_qps=0.0, duration_s=2.0,
                 max_outstanding=64, seed=1) -> dict:
        """Open-loop load from `threads` host threads (C++), returns stats."""
        L = np.ascontiguousarray(layers, np.float32)
        C = np.ascontiguousarray(cuboids, np.float32)
        K = np.ascontiguousarray(kinds, np.int32)
        Nn = np.ascontiguousarray(Ns, np.int32)
        A = np.ascontiguousarray(axes, np.int32)
        cfg = LoadConfig(threads, max_outstanding, rate_qps, duration_s, seed)
        st = ServerStats()
        check(_lib.lib().drw_server_run_load(self._h, ctypes.byref(cfg), L.ctypes.data, C.ctypes.data, K.ctypes.data,
                                             Nn.ctypes.data, A.ctypes.data, len(K), ctypes.byref(st)),
              "drw_server_run_load")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def capacitance(self, geometry: dict, master: int, walks=100000, threads=16, concurrent=64, N=23,
                    stop_dist=None, gauss_frac=0.5, max_cube=None, seed=1):
        """FRW capacitance row C[master][:] of a z-layered structure.

        geometry: window (6,), layer_z (n,), layer_eps (n,), boxes (m, 6), ids (m,),
        optionally media_boxes (k, 6) and media_eps (k,) (dielectric boxes)."""
        lz = np.ascontiguousarray(geometry["layer_z"], np.float64)
        le = np.ascontiguousarray(geometry["layer_eps"], np.float64)
        bx = np.ascontiguousarray(geometry["boxes"], np.float64)
        ids = np.ascontiguousarray(geometry["ids"], np.int32)
        mb = np.ascontiguousarray(geometry.get("media_boxes", np.zeros((0, 6))), np.float64).reshape(-1, 6)
        me = np.ascontiguousarray(geometry.get("media_eps", np.zeros(0)), np.float64)
        win = np.asarray(geometry["window"], np.float64)
        scale = float(np.min(bx[:, 3:] - bx[:, :3]))
        g = Geometry((ctypes.c_double * 6)(*win), len(lz), lz.ctypes.data, le.ctypes.data, len(ids), bx.ctypes.data,
                     ids.ctypes.data, len(me), mb.ctypes.data if len(me) else None, me.ctypes.data if len(me) else None)
        wc = WalkConfig(walks, threads, concurrent, N, stop_dist or 1e-3 * scale, gauss_frac,
                        max_cube or float(np.max(win[3:] - win[:3])), seed)
        nc = int(ids.max()) + 1
        cap = np.zeros(nc)
        err = np.zeros(nc)
        st = np.zeros(10)
        check(_lib.lib().drw_walk_capacitance(self._h, ctypes.byref(g), master, ctypes.byref(wc), cap.ctypes.data,
                                              err.ctypes.data, st.ctypes.data), "drw_walk_capacitance")
        return cap, err, {"walks": st[0], "hops_per_walk": st[1], "gauss_area": st[2], "gauss_margin": st[3],
                          "window_walks": st[4], "stalled_walks": st[5], "self_direct": st[6],
                          "self_direct_stderr": st[7], "mean_weight": st[8], "mean_weight_stderr": st[9]}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().drw_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
