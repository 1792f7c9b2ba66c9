"""Mixed walker batches: many transition queries of both kinds and mixed
resolution at once, each with its own Philox counter.

N = 23 queries go to the networks, N = 16 / 21 queries to the FDM path
(batched PCG solve + flat inverse-CDF over the 6 N^2 boundary tiles), one
launch sequence per (kind, N) group on one stream; a query's result depends
only on its description, axis, N and counter, never on the batch it rides in.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from . import structures as S
from .dnnsolver import DNNSolverGrad, DNNSolverPoisson

GROUP_NS = (16, 21, 23)


def make_queries(n: int, seed: int = 0, pool: int = 4096):
    """n queries: N uniform over {16, 21, 23}, kind alternating Poisson /
    gradient, axis uniform; descriptions drawn from a per-N pool of `pool`
    structures (layered_boxes at that N); counters 1..n."""
    rng = np.random.default_rng(seed)
    Ns = rng.choice(np.asarray(GROUP_NS, np.int32), size=n)
    kinds = (np.arange(n) % 2).astype(np.int32)
    axes = rng.integers(0, 3, n).astype(np.int32)
    L = np.empty((n, 15, 2), np.float32)
    C = np.empty((n, 15, 7), np.float32)
    for k, N in enumerate(GROUP_NS):
        Lp, Cp, _ = S.layered_boxes(pool, N, seed=seed * 7919 + k)
        sel = np.nonzero(Ns == N)[0]
        j = rng.integers(0, pool, sel.size)
        L[sel] = Lp[j]
        C[sel] = Cp[j]
    qid = np.arange(1, n + 1, dtype=np.uint64)
    return {"N": Ns, "kind": kinds, "axis": axes, "layers": L, "cuboids": C, "qid": qid}


class MixedRunner:
    """One solver per (kind, N) group: networks for N = 23 (`models_txt`),
    FDM at `fdm_tol` otherwise."""

    def __init__(self, models_txt: str, seed: int = 0x5EED, fdm_tol: float = 1e-6, device: int = 0):
        self.solvers = {}
        for kind in (0, 1):
            for N in GROUP_NS:
                s = (DNNSolverPoisson if kind == 0 else DNNSolverGrad)(device)
                if N == 23:
                    s.loadModels(models_txt=models_txt)
                else:
                    s.use_fdm(N, fdm_tol)
                s.set_seed(seed, 0)
                self.solvers[(kind, N)] = s

    def run(self, q: dict, subset=None):
        """Runs the queries (all, or the indices in `subset`); returns
        (out float32 [n, 5] with Poisson rows in columns 0..2, wall seconds
        per group (each group synchronised), total seconds)."""
        import torch
        idx_all = np.arange(q["N"].size) if subset is None else np.asarray(subset)
        out = np.full((q["N"].size, 5), np.nan, np.float32)
        plan = []
        for (kind, N), s in self.solvers.items():
            sel = idx_all[(q["kind"][idx_all] == kind) & (q["N"][idx_all] == N)]
            if sel.size == 0:
                continue
            dL = torch.from_numpy(np.ascontiguousarray(q["layers"][sel])).cuda()
            dC = torch.from_numpy(np.ascontiguousarray(q["cuboids"][sel])).cuda()
            dA = torch.from_numpy(np.ascontiguousarray(q["axis"][sel])).cuda()
            dQ = torch.from_numpy(np.ascontiguousarray(q["qid"][sel]).view(np.int64)).cuda()
            w = 3 if kind == 0 else 5
            dO = torch.empty((sel.size, w), device="cuda")
            plan.append((kind, N, s, sel, dL, dC, dA, dQ, dO))
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        group_s = {}
        t0 = time.perf_counter()
        for kind, N, s, sel, dL, dC, dA, dQ, dO in plan:
            tg = time.perf_counter()
            _lib.check(_lib.lib().drw_solver_sample_ids(s.handle, sel.size, dL.data_ptr(), dC.data_ptr(),
                                                         dA.data_ptr() if kind == 1 else None, dQ.data_ptr(),
                                                         dO.data_ptr(), stream.cuda_stream), "sample_ids")
            torch.cuda.synchronize()
            group_s[f"{'poisson' if kind == 0 else 'gradient'}_N{N}"] = {"queries": int(sel.size),
                                                                          "seconds": time.perf_counter() - tg}
        total = time.perf_counter() - t0
        for kind, N, s, sel, dL, dC, dA, dQ, dO in plan:
            out[sel, :dO.shape[1]] = dO.cpu().numpy()
        return out, group_s, total
