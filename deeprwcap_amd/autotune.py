"""Launch-shape and batch-shape autotuner.

Two searches, both timed on the device with the solver's stage timers (HIP
events on the launch stream):

* launch shapes: threads per workgroup of each face-predictor kernel (256 or
  512 per (query, face); a wider workgroup shares one query's LDS among more
  waves, which pays where the kernel is latency-bound: small batches, and the
  gradient kernels, whose LDS allows only two workgroups per CU), and for
  Poisson queries the fused query kernel at 256 or 512 threads against the
  separate selector, head and predictor kernels (shape 1).  Every kernel is
  timed at each shape over a ladder of batch sizes (the Poisson query as a
  whole: the sum of its stages) and the winners become the library's
  per-kernel rules (drw_set_launch_shape: from `min_items` workgroups on,
  use `threads`); a batch size switches shape only when another one is
  faster by more than 2 %;
* batch shape: with those shapes, one batch of every route the server uses
  (network Poisson / gradient queries at N = 23, FDM Poisson / gradient
  queries at N = 16, 21, 23) over a ladder of batch sizes, and the server's
  `max_batch` = the largest batch whose slowest route still fits in
  `latency_ms / slots` (a request can wait behind the batches in the other
  slots before its own runs), so that p99 latency stays inside the budget
  while batches stay as large as possible.

    python -m deeprwcap_amd.autotune --models models/models.txt --out tuning.json

`apply(tuning)` sets the launch shapes in this process;
`deeprwcap_amd.serving.Server.from_tuning(path, ...)` applies them and builds
a server with the tuned batch shape.
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
SHAPE_LADDER = (64, 128, 256, 512, 768, 1024, 1536, 2048, 4096, 8192)
SHAPES = (256, 512)
# drw_shape ids (include/dnnsolver_c.h)
SHAPE_KERNELS = {"face_poisson": 0, "face_grad1": 1, "face_grad2": 2, "poisson_fused": 3}
FUSED_SHAPES = (1, 256, 512)  # 1: the separate kernels
SWITCH_MARGIN = 0.02


def set_shape_rules(which, rules):
    """Replace kernel `which`'s rules by [(min_items, threads), ...]."""
    lib = _lib.lib()
    _lib.check(lib.drw_set_launch_shape(which, -1, int(rules[0][1]) if rules else 0), "set_launch_shape")
    for mn, th in rules[1:]:
        _lib.check(lib.drw_set_launch_shape(which, int(mn), int(th)), "set_launch_shape")


def apply(tuning):
    """Set the launch shapes of a tuning result (dict or JSON path) in this process."""
    if isinstance(tuning, str):
        with open(tuning) as f:
            tuning = json.load(f)
    for name, rules in tuning.get("launch_shapes", {}).items():
        set_shape_rules(SHAPE_KERNELS[name], rules)


def rules_from_timings(times, margin=SWITCH_MARGIN):
    """times: {batch: {threads: ms}} -> [(min_items, threads), ...] with the
    first rule at 0.  Walking up the batch sizes, the shape changes only where
    the other one is faster by more than `margin`; the rule starts at that
    batch size."""
    rules, cur = [], None
    for B in sorted(times):
        t = times[B]
        best = min(t, key=t.get)
        if cur is None:
            cur = best
            rules.append((0, cur))
        elif best != cur and t[best] < (1.0 - margin) * t[cur]:
            cur = best
            rules.append((int(B), cur))
    return rules


def _stage_ms(solver, run, stages, reps, torch):
    """Mean milliseconds per launch of one stage, or the sum over a tuple of
    stages (per launch of the first that ran)."""
    import ctypes
    lib = _lib.lib()
    ms = (ctypes.c_double * 8)()
    cnt = (ctypes.c_int64 * 8)()
    for _ in range(3):
        run()
    torch.cuda.synchronize()
    _lib.check(lib.drw_solver_stage_times(solver.handle, ms, cnt, 1), "stage_times")
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    _lib.check(lib.drw_solver_stage_times(solver.handle, ms, cnt, 1), "stage_times")
    stages = (stages,) if isinstance(stages, int) else stages
    n = max([cnt[k] for k in stages] + [1])
    return sum(ms[k] for k in stages) / n


def tune_shapes(models_txt, ladder=SHAPE_LADDER, device=0, log=print):
    """Face-kernel launch shapes: {"launch_shapes": {kernel: rules}, "shape_ms": ...}."""
    import torch
    torch.cuda.set_device(device)
    lib = _lib.lib()
    Bmax = max(ladder)
    L, C, _ = S.layered_boxes(Bmax, 23, seed=23)
    A = np.random.default_rng(23).integers(0, 3, Bmax).astype(np.int32)
    dL, dC, dA = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (L, C, A))
    stream = torch.cuda.current_stream().cuda_stream
    times = {k: {} for k in SHAPE_KERNELS}
    for kind in (0, 1):
        s = (DNNSolverPoisson if kind == 0 else DNNSolverGrad)(device)
        s.loadModels(models_txt=models_txt)
        _lib.check(lib.drw_solver_set_timing(s.handle, 1), "set_timing")
        out = torch.empty((Bmax, 3 if kind == 0 else 5), device="cuda")
        for B in ladder:
            def run():
                _lib.check(lib.drw_solver_sample_device(s.handle, B, dL.data_ptr(), dC.data_ptr(),
                                                        dA.data_ptr() if kind else None, out.data_ptr(), stream),
                           "sample_device")
            reps = max(10, min(60, int(4e5 / B)))
            if kind == 0:
                # separate kernels (the face kernel at each shape), then the
                # fused query kernel at each shape; whole-query times
                set_shape_rules(3, [(0, 1)])
                t = {}
                for th in SHAPES:
                    set_shape_rules(0, [(0, th)])
                    t[th] = _stage_ms(s, run, 2, reps, torch)
                times["face_poisson"][B] = t
                set_shape_rules(0, [(0, min(t, key=t.get))])
                tf = {1: _stage_ms(s, run, (0, 1, 2), reps, torch)}
                for th in SHAPES:
                    set_shape_rules(3, [(0, th)])
                    tf[th] = _stage_ms(s, run, 7, reps, torch)
                times["poisson_fused"][B] = tf
            else:
                # the stage runs Gradient1 then Gradient2: vary one kernel at a time
                base = {}
                for k1, k2 in ((256, 256), (512, 256), (256, 512)):
                    set_shape_rules(1, [(0, k1)])
                    set_shape_rules(2, [(0, k2)])
                    base[(k1, k2)] = _stage_ms(s, run, 2, reps, torch)
                t0 = base[(256, 256)]
                times["face_grad1"][B] = {256: t0, 512: base[(512, 256)]}
                times["face_grad2"][B] = {256: t0, 512: base[(256, 512)]}
            log(f"shape B={B}: " + ", ".join(f"{k} " + "/".join(f"{v[B][th]:.4f}" for th in sorted(v[B]))
                                               for k, v in times.items() if B in v)
                + " ms (256/512 threads; poisson_fused: separate/256/512)")
        del s
    rules = {k: rules_from_timings(v) for k, v in times.items()}
    for k, r in rules.items():
        set_shape_rules(SHAPE_KERNELS[k], r)
        log(f"{k}: " + ", ".join(f"{th} threads from {mn} workgroups" for mn, th in r))
    return {"launch_shapes": {k: [list(x) for x in r] for k, r in rules.items()},
            "shape_ms": {k: {str(B): {str(th): ms for th, ms in t.items()} for B, t in v.items()}
                         for k, v in times.items()}}


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


def tune(models_txt, latency_ms=1.0, slots=4, ladder=LADDER, fdm_tol=1e-6, reps=7, device=0, log=print,
         shapes=True):
    import torch
    torch.cuda.set_device(device)
    shape_res = tune_shapes(models_txt, device=device, log=log) if shapes and models_txt else {}
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
           "batch_ms": {k: {str(B): v for B, v in d.items()} for k, d in per.items()}, **shape_res}
    log(f"max_batch = {max_batch} (slowest network route within {budget:.3f} ms per batch)")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--models", default=None, help="models.txt (omit: FDM routes only)")
    ap.add_argument("--latency-ms", type=float, default=1.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--out", default="tuning.json")
    ap.add_argument("--no-shapes", action="store_true", help="keep the library's launch shapes")
    args = ap.parse_args()
    res = tune(args.models, latency_ms=args.latency_ms, slots=args.slots, shapes=not args.no_shapes)
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
