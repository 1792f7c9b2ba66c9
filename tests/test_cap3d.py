"""cap3d reader on the reference's own small test structures
(testcases/cap3d/case4-6.cap3d and their field-solver results ref/case*.dspf,
kept as fixtures under tests/golden/cap3d)."""
import os

import numpy as np
import pytest

from deeprwcap_amd import cap3d

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cap3d")
CASES = {4: ("NoLayerALRDLM4", 8.210106e-15), 5: ("NoLayerTMpoly", 1.419204e-14), 6: ("NoLayerTMM1", 9.584033e-15)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_parse_case(case):
    s = cap3d.read(os.path.join(HERE, f"case{case}.cap3d"))
    net, total = cap3d.read_reference_total(open(os.path.join(HERE, f"case{case}.dspf")).read())
    assert (net, total) == CASES[case]
    assert s["masters"] == [net] and net in s["names"]
    # planar layers bottom-up, first one open below
    assert s["layer_z"][0] < -1e20 and np.all(np.diff(s["layer_z"]) > 0)
    assert len(s["layer_z"]) == len(s["layer_eps"]) and (s["layer_eps"] >= 1.0).all()
    # boxes are well formed, every conductor has at least one
    for b in (s["boxes"], s["media_boxes"]):
        assert (b[:, 3:] > b[:, :3]).all()
    assert set(s["ids"]) == set(range(len(s["names"])))
    assert len(s["media_eps"]) == len(s["media_boxes"]) > 0
    # everything lies inside the window
    w = s["window"]
    assert (s["boxes"][:, :3] >= w[:3] - 1e-9).all() and (s["boxes"][:, 3:] <= w[3:] + 1e-9).all()


def test_block_fields():
    text = """<cap3d><window> v1(0,0,0) v2(10,10,10) </window>
<plate_medium> diel 2.0 z_top 1.0 </plate_medium>
<plate_medium> diel 3.0 z_top 10.0 </plate_medium>
<medium> <block> basepoint(1,1,1) v1(2,0,0) v2(0,3,0) hvector(0,0,1) </block> diel 5.0 </medium>
<conductor> name A <block> basepoint(5,5,5) v1(1,0,0) v2(0,1,0) hvector(0,0,-1) </block> </conductor>
<task><capacitance> A </capacitance></task></cap3d>"""
    s = cap3d.parse(text)
    np.testing.assert_array_equal(s["boxes"][0], [5, 5, 4, 6, 6, 5])  # negative edge vector
    np.testing.assert_array_equal(s["media_boxes"][0], [1, 1, 1, 3, 4, 2])
    assert s["media_eps"].tolist() == [5.0] and s["layer_eps"].tolist() == [2.0, 3.0]
    assert s["layer_z"][1] == 1.0 and s["masters"] == ["A"]


def test_rejects_skewed_block():
    text = """<cap3d><window> v1(0,0,0) v2(1,1,1) </window>
<conductor> name A <block> basepoint(0,0,0) v1(1,1,0) v2(0,1,0) hvector(0,0,1) </block> </conductor></cap3d>"""
    with pytest.raises(ValueError):
        cap3d.parse(text)


def test_net_names():
    assert cap3d.net_of("7__signalGND") == "signalGND"
    assert cap3d.net_of("ln_13") == "ln_13" and cap3d.net_of("a__b") == "a__b" and cap3d.net_of("3__") == "3__"
    s = cap3d.read(os.path.join(HERE, "case9.cap3d.gz"))
    order, ids = cap3d.nets(s)
    # 47 conductor sections, 18 nets (ln_13 and others listed in several sections)
    assert len(s["names"]) == 47 and len(order) == 18 and s["names"].count("ln_13") > 1
    assert ids.shape == s["ids"].shape and set(ids) == set(range(18))
    for k, name in enumerate(s["names"]):
        assert (ids[s["ids"] == k] == order.index(cap3d.net_of(name))).all()


@pytest.mark.parametrize("case", range(1, 11))
def test_reference_nets_exist(case):
    """Every net a reference row names is a net of the structure."""
    p = os.path.join(HERE, f"case{case}.cap3d")
    s = cap3d.read(p if os.path.exists(p) else p + ".gz")
    ref = cap3d.read_reference(os.path.join(HERE, f"case{case}.dspf" if case < 10 else "case10.spice"))
    order, _ = cap3d.nets(s)
    assert {n for row in ref.values() for n in row} <= set(order)
    assert set(ref) <= set(order)


def test_reference_matrix_dspf():
    ref = cap3d.read_reference(os.path.join(HERE, "case7.dspf"))
    assert list(ref) == ["1", "2", "4", "5", "6"]
    assert ref["1"]["1"] == 9.176692675804292e-16
    # a coupling in both rows is their mean, in both rows
    m = (1.849878959329723e-17 + 1.8502064456341546e-17) / 2
    assert ref["1"]["2"] == ref["2"]["1"] == m
    assert ref["1"]["3"] == 1.6741187825602342e-17 and "3" not in ref
    assert set(ref["1"]) == {"1", "2", "3", "4", "5", "6", "GROUND", "signalGND"}
    # the field solver's totals are the sums of their rows' couplings
    for k, row in ref.items():
        assert sum(v for n, v in row.items() if n != k) == pytest.approx(row[k], rel=2e-3)
    text = "*|NET A 1e-15\nC1 A 0 2e-16\nC2 A B 3e-16\n"
    assert cap3d.read_reference_matrix(text) == {"A": {"A": 1e-15, "GROUND": 2e-16, "B": 3e-16}}


def test_reference_matrix_spice():
    ref = cap3d.read_reference(os.path.join(HERE, "case10.spice"))
    assert list(ref) == ["G"] and len(ref["G"]) == 8
    assert ref["G"]["G"] == pytest.approx(1.241e-15, rel=1e-12)
    assert ref["G"]["S"] == pytest.approx(0.5735e-15, rel=1e-12)
    assert ref["G"]["ln_14"] == 2.44e-18
    assert cap3d.read_reference_spice("** X 2pF * (total)\nCp1 X 0 1.5fF\n") == \
        {"X": {"X": pytest.approx(2e-12), "GROUND": pytest.approx(1.5e-15)}}


def test_matrix_error():
    ref = cap3d.read_reference(os.path.join(HERE, "case7.dspf"))
    err, self_err = cap3d.matrix_error(ref, ref)
    assert err == 0 and self_err == {k: 0 for k in ref}
    total = sum(abs(v) for row in ref.values() for v in row.values())
    # every entry 2 % high
    hi = {m: {n: 1.02 * v for n, v in row.items()} for m, row in ref.items()}
    err, self_err = cap3d.matrix_error(hi, ref)
    assert err == pytest.approx(0.02) and all(v == pytest.approx(0.02) for v in self_err.values())
    # a missing coupling counts as zero
    drop = {m: dict(row) for m, row in ref.items()}
    del drop["1"]["3"]
    assert cap3d.matrix_error(drop, ref)[0] == pytest.approx(ref["1"]["3"] / total)
    # rows the matrix does not hold are left out of both sums
    one = {"1": {n: 1.1 * v for n, v in ref["1"].items()}}
    err, self_err = cap3d.matrix_error(one, ref)
    assert err == pytest.approx(0.1) and list(self_err) == ["1"]
    # extra entries of the matrix are not scored
    more = {m: dict(row, extra=1.0) for m, row in ref.items()}
    assert cap3d.matrix_error(more, ref)[0] == 0
    assert np.isnan(cap3d.matrix_error(None, ref)[0])
