"""Batch-shape autotuner for the serving engine.

Measures the device time of one batch on every route the server uses
(network Poisson / gradient queries at N = 23, FDM Poisson / gradient queries
at N = 16, 21, 23) over a ladder of batch sizes, and picks the server's
`max_batch`: the largest batch whose slowest route still fits in
`latency_ms / slots` (a request can wait behind the batches in the other
slots before its own runs), so that p99 latency stays inside the budget while
batches stay as large as possible for throughput.

    python -m deeprwcap_amd.autotune --models models/models.txt --out tuning.json

`deeprwcap_amd.serving.Server.from_tuning(path, ...)` builds a server from
the result.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib
from . import structures as S
from .dnnsolver import DNNSolverGrad, DNNSolverPoisson

LADDER = (64, 128, 256, 512, 1024, 2048, 4096)


def _batch_ms(solver, kind, B, L, C, A, reps, torch):
    dL = torch.from_numpy(np.ascontiguousarray(L[:B])).cuda()
    dC = torch.from_numpy(np.ascontiguousarray(C[:B])).cuda()
    dA = torch.from_numpy(np.ascontiguousarray(A[:B])).cuda()
    out = torch.empty((B, 3 if kind == 0 else 5), device="cuda")
    stream = torch.cuda.current_stream()

    def run():
        _lib.check(_lib.lib().drw_solver_sample_device(solver.handle, B, dL.data_ptr(), dC.data_ptr(),
                                                        dA.data_ptr() if kind == 1 else None, out.data_ptr(),
                                                        stream.cuda_stream), "sample_device")

    run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return 1e3 * float(np.median(ts))


def tune(models_txt, latency_ms=1.0, slots=4, ladder=LADDER, fdm_tol=1e-6, reps=7, device=0, log=print):
    import torch
    torch.cuda.set_device(device)
    routes = {}
    for kind in (0, 1):
        for N in (16, 21, 23):
            for backend in (("nn", "fdm") if N == 23 else ("fdm",)):
                if backend == "nn" and not models_txt:
                    continue
                s = (DNNSolverPoisson if kind == 0 else DNNSolverGrad)(device)
                if backend == "nn":
                    s.loadModels(models_txt=models_txt)
                else:
                    s.use_fdm(N, fdm_tol)
                routes[(kind, N, backend)] = s
    Lmax = max(ladder)
    per = {}
    for (kind, N, backend), s in routes.items():
        L, C, _ = S.layered_boxes(Lmax, N, seed=17 + N)
        A = np.random.default_rng(N).integers(0, 3, Lmax).astype(np.int32)
        key = f"{'poisson' if kind == 0 else 'gradient'}_{backend}_N{N}"
        per[key] = {}
        for B in ladder:
            per[key][B] = _batch_ms(s, kind, B, L, C, A, reps, torch)
            log(f"{key} B={B}: {per[key][B]:.3f} ms ({B / per[key][B] * 1e3:,.0f} queries/s)")
    budget = latency_ms / slots
    # the network routes decide the batch when models are loaded (N = 23 is the
    # serving path); FDM routes only bound it when they are the only ones
    keys = [k for k in per if "_nn_" in k] or list(per)
    fit = [B for B in ladder if max(per[k][B] for k in keys) <= budget]
    max_batch = max(fit) if fit else min(ladder)
    res = {"max_batch": int(max_batch), "max_inflight": slots, "latency_budget_ms": latency_ms,
           "batch_ms": {k: {str(B): v for B, v in d.items()} for k, d in per.items()}}
    log(f"max_batch = {max_batch} (slowest network route within {budget:.3f} ms per batch)")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--models", default=None, help="models.txt (omit: FDM routes only)")
    ap.add_argument("--latency-ms", type=float, default=1.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--out", default="tuning.json")
    args = ap.parse_args()
    res = tune(args.models, latency_ms=args.latency_ms, slots=args.slots)
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
