"""GPU tests of the serving engine's resident loop (drw_server_config.resident):
a persistent kernel takes N = 23 network queries from a pinned request ring
and writes each result back with no host step per query.  Every result must
be bit-identical to the batched route's for the same Philox counter, for both
kinds, both workgroup sizes, alone and in a concurrent mix, across kernel
restarts (idle end, FDM batches of other N in between)."""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "models", "models.txt")
WAIT_MS = 20000  # a host-side wait that ends in an error instead of hanging


def _requests(n, seed, Ns=(23,)):
    from deeprwcap_amd import structures as S
    rng = np.random.default_rng(seed)
    N = rng.choice(np.array(Ns), n)
    L, C, _ = S.layered_boxes(n, N, seed=seed)
    kinds = rng.integers(0, 2, n)
    axes = rng.integers(0, 3, n)
    return [(int(kinds[i]), int(N[i]), int(axes[i]), L[i], C[i], 7000 + 3 * i) for i in range(n)]


def _serve_many(sv, reqs):
    t = sv.submit_many([r[0] for r in reqs], [r[1] for r in reqs], [r[2] for r in reqs],
                       np.stack([r[3] for r in reqs]), np.stack([r[4] for r in reqs]),
                       np.array([r[5] for r in reqs], np.uint64))
    return sv.wait_many(t, timeout_ms=WAIT_MS)


def _batched(reqs):
    from deeprwcap_amd.serving import Server
    sv = Server(0, MODELS, max_batch=64, max_inflight=4)
    out = _serve_many(sv, reqs)
    sv.close()
    return out


@pytest.mark.parametrize("threads", [512, 256])
def test_resident_matches_batched(threads):
    """N = 23 queries of both kinds (shipped weights): the resident kernel's
    results equal the batched route's, submitted all at once and one by one."""
    from deeprwcap_amd.serving import Server
    reqs = _requests(384, 3)
    ref = _batched(reqs)
    print(f"resident {threads} threads: {len(reqs)} queries", flush=True)
    sv = Server(0, MODELS, resident=True, resident_threads=threads)
    t0 = time.perf_counter()
    many = _serve_many(sv, reqs)
    t1 = time.perf_counter()
    single = [sv.wait(sv.submit(*r), r[0], timeout_ms=WAIT_MS) for r in reqs[:48]]
    st = sv.stats()
    print(f"  resident counters {sv.resident_stats()}", flush=True)
    sv.close()
    print(f"  all at once {1e3 * (t1 - t0):.2f} ms; stats {st}", flush=True)
    assert np.isfinite(ref).all()
    for i, r in enumerate(reqs):
        w = 3 if r[0] == 0 else 5
        np.testing.assert_array_equal(many[i, :w], ref[i, :w], err_msg=f"query {i} {r[:3]}")
    for i, s in enumerate(single):
        np.testing.assert_array_equal(s, ref[i, :len(s)], err_msg=f"query {i} alone")
    assert st["completed"] == len(reqs) + 48


def test_resident_concurrent_and_restart():
    """Eight submitting threads at once, then an idle gap longer than the
    kernel's idle limit (the kernel ends), then more queries (the dispatcher
    starts it again): every result equals the batched one."""
    from deeprwcap_amd.serving import Server
    reqs = _requests(256, 5)
    ref = _batched(reqs)
    sv = Server(0, MODELS, resident=True, resident_idle_ms=2.0)
    with ThreadPoolExecutor(8) as ex:
        tickets = list(ex.map(lambda r: sv.submit(*r), reqs[:192]))
    mix = [sv.wait(t, r[0], timeout_ms=WAIT_MS) for t, r in zip(tickets, reqs[:192])]
    time.sleep(0.05)
    late = _serve_many(sv, reqs[192:])
    sv.close()
    for i, m in enumerate(mix):
        np.testing.assert_array_equal(m, ref[i, :len(m)], err_msg=f"query {i}")
    for j, r in enumerate(reqs[192:]):
        w = 3 if r[0] == 0 else 5
        np.testing.assert_array_equal(late[j, :w], ref[192 + j, :w], err_msg=f"late query {j}")


def test_resident_with_fdm_route():
    """Mixed N {16, 21, 23}: the FDM-route batches (N = 16, 21) stop the
    resident kernel and it restarts after them; all results equal the
    batched server's."""
    from deeprwcap_amd.serving import Server
    reqs = _requests(240, 9, Ns=(16, 21, 23))
    ref = _batched(reqs)
    sv = Server(0, MODELS, resident=True, max_batch=64)
    got = _serve_many(sv, reqs)
    st = sv.stats()
    sv.close()
    print(f"mixed N through the resident server: {st}", flush=True)
    for i, r in enumerate(reqs):
        w = 3 if r[0] == 0 else 5
        np.testing.assert_array_equal(got[i, :w], ref[i, :w], err_msg=f"query {i} {r[:3]}")


@pytest.mark.parametrize("rate", [1e5, 1e6])
def test_resident_load_probes(rate):
    """Config-5 serving load on the resident loop: 64 host threads issue
    Poisson-arrival N = 23 queries of both kinds at the offered rate; probe
    queries submitted into that mix equal their results alone."""
    import threading
    from deeprwcap_amd import structures as S
    from deeprwcap_amd.serving import Server
    pool = 4096
    rng = np.random.default_rng(13)
    Ns = np.full(pool, 23, np.int32)
    L, C, _ = S.layered_boxes(pool, 23, seed=23)
    kinds = rng.integers(0, 2, pool).astype(np.int32)
    axes = rng.integers(0, 3, pool).astype(np.int32)
    probes = _requests(96, 33)
    sv = Server(0, MODELS, resident=True)
    sv.run_load(L, C, kinds, Ns, axes, threads=64, rate_qps=0, duration_s=0.2)  # warm-up
    res = {}
    th = threading.Thread(target=lambda: res.update(sv.run_load(L, C, kinds, Ns, axes, threads=64, rate_qps=rate,
                                                                duration_s=1.5, max_outstanding=64)))
    print(f"resident load {rate:.0e}/s with {len(probes)} probes", flush=True)
    th.start()
    time.sleep(0.3)
    mix = _serve_many(sv, probes)
    th.join()
    alone = [sv.wait(sv.submit(*r), r[0], timeout_ms=WAIT_MS) for r in probes]
    rs = sv.resident_stats()
    sv.close()
    print(f"offered {rate:.0e}/s: achieved {res['qps']:.4g}/s, p50 {res['p50_ms']:.3f} ms, "
          f"p99 {res['p99_ms']:.3f} ms; resident {rs}", flush=True)
    assert abs(res["qps"] / rate - 1) < 0.1
    assert res["p99_ms"] >= res["p50_ms"] > 0
    for i, (r, a) in enumerate(zip(probes, alone)):
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(mix[i, :len(a)], a, err_msg=f"probe {i} {r[:3]}")


def test_resident_closed_loop_rate():
    """Closed-loop rate of the resident loop: 16,384 N = 23 queries submitted
    in one call and waited for (mixed kinds, Poisson only, gradient only;
    both workgroup sizes), against the batched route (reported)."""
    from deeprwcap_amd import structures as S
    from deeprwcap_amd.serving import Server
    n = 16384
    rng = np.random.default_rng(17)
    L, C, _ = S.layered_boxes(n, 23, seed=29)
    mixed = rng.integers(0, 2, n).astype(np.int32)
    axes = rng.integers(0, 3, n).astype(np.int32)
    for resident, threads in ((True, 512), (True, 256), (False, 0)):
        sv = Server(0, MODELS, resident=resident, resident_threads=threads or 512)
        for name, kinds in (("mixed", mixed), ("poisson", np.zeros(n, np.int32)), ("gradient", np.ones(n, np.int32))):
            for rep in range(2):
                t0 = time.perf_counter()
                t = sv.submit_many(kinds, np.full(n, 23, np.int32), axes, L, C,
                                   np.arange(rep * n + 1, (rep + 1) * n + 1, dtype=np.uint64))
                t1 = time.perf_counter()
                out = sv.wait_many(t, timeout_ms=WAIT_MS)
                t2 = time.perf_counter()
                assert np.isfinite(out[:, :3]).all()
            print(f"{'resident ' + str(threads) if resident else 'batched'} {name}: submit {1e3 * (t1 - t0):.2f} ms, "
                  f"total {1e3 * (t2 - t0):.2f} ms, {n / (t2 - t0):,.0f} queries/s", flush=True)
        if resident:
            print(f"  resident counters {sv.resident_stats()}", flush=True)
        sv.close()


def test_resident_network_failure_goes_to_fdm(tmp_path):
    """Models whose selector head biases are NaN make every resident draw
    fail: the kernel answers with status 1, the dispatcher re-runs each query
    on the FDM route at N = 23 with the same counter, and every result equals
    a model-less server's FDM result for that counter (both kinds)."""
    from deeprwcap_amd import weights as W
    from deeprwcap_amd.serving import Server
    states = {n: W.read_drwnet(os.path.join(ROOT, "models", f"{n}.drwnet")) for n in W.MODEL_ORDER}
    for sel in ("PoissonSelector", "GradientSelectorWeight"):
        last = max(k for k in states[sel] if k.startswith("head.") and k.endswith(".bias"))
        states[sel][last] = np.full_like(states[sel][last], np.nan)
    broken = W.export_model_dir(str(tmp_path / "broken"), states)
    reqs = _requests(96, 41)
    fdm = Server(0, None, max_batch=64)
    want = _serve_many(fdm, reqs)
    fdm.close()
    print(f"resident network failures: {len(reqs)} queries on NaN-bias selectors", flush=True)
    sv = Server(0, broken, resident=True)
    got = _serve_many(sv, reqs)
    rs = sv.resident_stats()
    sv.close()
    print(f"  resident counters {rs}", flush=True)
    for i, r in enumerate(reqs):
        w = 3 if r[0] == 0 else 5
        assert np.isfinite(got[i, :w]).all(), f"query {i} {r[:3]}"
        np.testing.assert_array_equal(got[i, :w], want[i, :w], err_msg=f"query {i} {r[:3]}")
