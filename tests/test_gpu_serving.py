"""GPU tests of the serving engine: each query's result is bit-identical
whether it runs alone or in a concurrent mix of kinds and resolutions, and
the latency statistics are reported."""
import os
from concurrent.futures import ThreadPoolExecutor

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


def _requests(n, seed):
    from deeprwcap_amd import structures as S
    rng = np.random.default_rng(seed)
    Ns = rng.choice([16, 21, 23], n)
    L, C, _ = S.layered_boxes(n, Ns, seed=seed)
    kinds = rng.integers(0, 2, n)
    axes = rng.integers(0, 3, n)
    return [(int(kinds[i]), int(Ns[i]), int(axes[i]), L[i], C[i], 5000 + i) for i in range(n)]


def test_alone_equals_mix(models_txt):
    from deeprwcap_amd.serving import Server
    reqs = _requests(160, 3)
    sv = Server(0, models_txt, max_batch=64, max_inflight=4, hbm_budget_mb=64)
    alone = []
    for k, n, a, L, C, q in reqs:
        alone.append(sv.wait(sv.submit(k, n, a, L, C, q), k))
    with ThreadPoolExecutor(8) as ex:
        tickets = list(ex.map(lambda r: sv.submit(r[0], r[1], r[2], r[3], r[4], r[5]), reqs))
    mix = [sv.wait(t, r[0]) for t, r in zip(tickets, reqs)]
    st = sv.stats()
    sv.close()
    for i, (a, m) in enumerate(zip(alone, mix)):
        np.testing.assert_array_equal(a, m, err_msg=f"request {i} {reqs[i][:3]}")
        assert np.isfinite(a).all()
    assert st["completed"] == 2 * len(reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    print("serving stats", st, flush=True)


def test_server_matches_solver(models_txt):
    """A served N=23 Poisson query equals the solver's batched result for the
    same Philox counter."""
    import torch
    from deeprwcap_amd import _lib
    from deeprwcap_amd.dnnsolver import DNNSolverPoisson
    from deeprwcap_amd.serving import Server
    from deeprwcap_amd import structures as S
    B = 48
    L, C, _ = S.layered_boxes(B, 23, seed=9)
    qid = np.arange(100, 100 + B, dtype=np.uint64)
    s = DNNSolverPoisson(0)
    s.loadModels(models_txt=models_txt)
    s.set_seed(77, 0)
    dL, dC, dq = torch.from_numpy(L).cuda(), torch.from_numpy(C).cuda(), torch.from_numpy(qid.view(np.int64)).cuda()
    out = torch.empty((B, 3), device="cuda")
    _lib.check(_lib.lib().drw_solver_sample_ids(s.handle, B, dL.data_ptr(), dC.data_ptr(), None, dq.data_ptr(),
                                                 out.data_ptr(), None), "sample_ids")
    torch.cuda.synchronize()
    ref = out.cpu().numpy()
    sv = Server(0, models_txt, max_batch=16, seed=77)
    got = [sv.wait(sv.submit(0, 23, 2, L[b], C[b], int(qid[b])), 0) for b in range(B)]
    sv.close()
    np.testing.assert_array_equal(np.stack(got), ref)


def test_submit_many_matches_single(models_txt):
    """The batch submit/wait calls return exactly what one-at-a-time calls
    return for the same query ids."""
    from deeprwcap_amd.serving import Server
    reqs = _requests(96, 5)
    sv = Server(0, models_txt, max_batch=32, max_inflight=3)
    single = [sv.wait(sv.submit(k, n, a, L, C, q), k) for k, n, a, L, C, q in reqs]
    kinds = [r[0] for r in reqs]
    t = sv.submit_many(kinds, [r[1] for r in reqs], [r[2] for r in reqs], np.stack([r[3] for r in reqs]),
                       np.stack([r[4] for r in reqs]), np.array([r[5] for r in reqs], np.uint64))
    many = sv.wait_many(t)
    sv.close()
    for i, (k, s) in enumerate(zip(kinds, single)):
        np.testing.assert_array_equal(many[i, :len(s)], s, err_msg=f"request {i} {reqs[i][:3]}")


@pytest.mark.parametrize("rate", [1e5, 1e6], ids=["1e5qps", "1e6qps"])
def test_load_with_probes_alone_equals_mix(rate):
    """Config-5 serving load: 64 host threads issue Poisson-arrival queries of
    mixed N {16, 21, 23} and both kinds at the offered rate (shipped weights);
    meanwhile probe queries with fixed counters are submitted into that mix.
    The achieved rate follows the offered one, p50/p99 latency are reported,
    and each probe's result in the mix equals its result alone."""
    import threading
    import time
    from deeprwcap_amd import structures as S
    from deeprwcap_amd.serving import Server
    pool = 4096
    rng = np.random.default_rng(11)
    Ns = rng.choice(np.array([16, 21, 23], np.int32), pool)
    L, C, _ = S.layered_boxes(pool, Ns, seed=21)
    kinds = rng.integers(0, 2, pool).astype(np.int32)
    axes = rng.integers(0, 3, pool).astype(np.int32)
    probes = _requests(96, 31)
    sv = Server(0, os.path.join(ROOT, "models", "models.txt"), max_batch=1024, max_inflight=4, hbm_budget_mb=512)
    sv.run_load(L, C, kinds, Ns, axes, threads=64, rate_qps=0, duration_s=0.2)  # warm-up
    res = {}
    th = threading.Thread(target=lambda: res.update(sv.run_load(L, C, kinds, Ns, axes, threads=64, rate_qps=rate,
                                                                duration_s=1.5, max_outstanding=64)))
    print(f"load {rate:.0e}/s with {len(probes)} probes", flush=True)
    th.start()
    time.sleep(0.3)
    qids = np.array([10 ** 9 + r[5] for r in probes], np.uint64)
    t = sv.submit_many([r[0] for r in probes], [r[1] for r in probes], [r[2] for r in probes],
                       np.stack([r[3] for r in probes]), np.stack([r[4] for r in probes]), qids)
    mix = sv.wait_many(t)
    th.join()
    alone = [sv.wait(sv.submit(r[0], r[1], r[2], r[3], r[4], int(q)), r[0]) for r, q in zip(probes, qids)]
    sv.close()
    print(f"offered {rate:.0e}/s: achieved {res['qps']:.4g}/s, p50 {res['p50_ms']:.3f} ms, p99 {res['p99_ms']:.3f} ms, "
          f"mean batch {res['mean_batch']:.1f}", flush=True)
    assert abs(res["qps"] / rate - 1) < 0.1
    assert res["p99_ms"] >= res["p50_ms"] > 0
    for i, (r, a) in enumerate(zip(probes, alone)):
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(mix[i, :len(a)], a, err_msg=f"probe {i} {r[:3]}")
