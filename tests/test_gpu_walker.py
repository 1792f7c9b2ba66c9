"""GPU tests of the FRW walker on the serving engine (the consumer of the
transition queries).

* Two independent FDM-only runs on two crossing wires (N = 16) agree within
  their statistical error, and the signs follow the physics (self term
  positive, couplings negative).
* Config 5 (2 x 8 crossing wires, 0.1 x 0.1 x 2 um at 0.2 um pitch, eps
  3.9 / 2.7 / 7.0), 64 walker threads at N = 23: the capacitance row from
  network queries (gradient first hop through GradientSelectorWeight /
  Gradient1/2Predictor, W sign eps_c weighting, then Poisson hops through
  PoissonSelector / PoissonPredictor, shipped weights) against the row from
  FDM-only queries.  Self-capacitance and total coupling agree within 1 %
  (the target), with a standard error of the difference near 0.2 % (8M
  network walks, 4M FDM-only walks).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_walker_two_wires_fdm_consistent():
    from deeprwcap_amd import structures as S
    from deeprwcap_amd.serving import Server
    geo = S.wire_crossing(n_wires=2)
    sv = Server(0, None, max_batch=2048, max_inflight=4, fdm_tol=1e-5)
    c1, e1, st1 = sv.capacitance(geo, 0, walks=6000, threads=8, concurrent=64, N=16, seed=1)
    c2, e2, st2 = sv.capacitance(geo, 0, walks=6000, threads=8, concurrent=64, N=16, seed=2)
    sv.close()
    print("run1", c1, e1, st1, flush=True)
    print("run2", c2, e2, st2, flush=True)
    assert c1[0] > 0 and (c1[1:] < 0).all()
    # independent runs agree within 4 combined standard errors
    assert abs(c1[0] - c2[0]) < 4 * np.hypot(e1[0], e2[0])
    assert 5 < st1["hops_per_walk"] < 500


def test_walker_config5_network_vs_fdm():
    from deeprwcap_amd import structures as S
    from deeprwcap_amd.serving import Server
    geo = S.wire_crossing(n_wires=8)
    rows = {}
    for name, mt, walks, seed in (("network", os.path.join(ROOT, "models", "models.txt"), 8_000_000, 11),
                                  ("fdm", None, 4_000_000, 12)):
        print(f"config 5 walker: {name}, {walks} walks, 64 threads, N=23", flush=True)
        sv = Server(0, mt, max_batch=4096, max_inflight=4, hbm_budget_mb=8192, fdm_tol=1e-6)
        cap, err, st = sv.capacitance(geo, 0, walks=walks, threads=64, concurrent=256, N=23, seed=seed)
        sv.close()
        print(f"{name}: C00 {cap[0]:.6g} +- {err[0]:.3g}, total coupling {cap[1:].sum():.6g}, "
              f"{st['hops_per_walk']:.1f} hops/walk", flush=True)
        assert cap[0] > 0 and (cap[1:] < 0).all()
        rows[name] = (cap, err)
    (a, ea), (b, eb) = rows["network"], rows["fdm"]
    d_self = (a[0] - b[0]) / b[0]
    s_self = np.hypot(ea[0], eb[0]) / b[0]
    cb = b[1:].sum()
    d_coup = (a[1:].sum() - cb) / cb
    s_coup = np.hypot(np.sqrt((ea[1:] ** 2).sum()), np.sqrt((eb[1:] ** 2).sum())) / abs(cb)
    print(f"network vs FDM-only: self {100 * d_self:+.3f}% (stderr {100 * s_self:.3f}%), total coupling "
          f"{100 * d_coup:+.3f}% (stderr {100 * s_coup:.3f}%)", flush=True)
    assert s_self < 0.3e-2 and s_coup < 0.3e-2
    assert abs(d_self) < 1e-2
    assert abs(d_coup) < 1e-2
