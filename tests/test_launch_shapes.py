"""Launch-shape rules of the C ABI (drw_set_launch_shape / drw_get_launch_shape,
host state only: no GPU needed) and the autotuner's rule selection."""
import json

import pytest


@pytest.fixture
def lib():
    from deeprwcap_amd import _lib
    L = _lib.lib()
    yield L
    for w in range(4):
        assert L.drw_set_launch_shape(w, -1, 0) == 0  # back to the built-in rules


def test_builtin_rules(lib):
    # the MI355X choice: 512 threads for small Poisson batches and for the
    # gradient kernels at every size, 256 for large Poisson batches
    assert [lib.drw_get_launch_shape(0, b) for b in (1, 64, 1023, 1024, 1536, 2048, 8192)] == \
        [512, 512, 512, 256, 512, 256, 256]
    for w in (1, 2):
        assert [lib.drw_get_launch_shape(w, b) for b in (1, 2048, 65536)] == [512] * 3
    # the fused Poisson query kernel below 4096 queries, the separate kernels
    # (shape 1) from there on
    assert [lib.drw_get_launch_shape(3, b) for b in (1, 1023, 1024, 1536, 4095, 4096, 8192)] == \
        [512, 512, 256, 512, 512, 1, 1]


def test_rule_semantics(lib):
    assert lib.drw_set_launch_shape(0, -1, 256) == 0          # one shape for every size
    assert {lib.drw_get_launch_shape(0, b) for b in (1, 100, 10 ** 6)} == {256}
    assert lib.drw_set_launch_shape(0, 100, 512) == 0          # from 100 workgroups on
    assert lib.drw_set_launch_shape(0, 1000, 256) == 0
    assert [lib.drw_get_launch_shape(0, b) for b in (99, 100, 999, 1000)] == [256, 512, 512, 256]
    assert lib.drw_set_launch_shape(0, 100, 256) == 0          # replaces the rule of 100
    assert lib.drw_get_launch_shape(0, 500) == 256
    assert lib.drw_set_launch_shape(0, 100, 0) == 0            # removes it
    assert lib.drw_set_launch_shape(0, 0, 0) == 0              # and the base rule: no rule left
    assert lib.drw_get_launch_shape(0, 500) == 256             # the default shape
    # the other kernels are untouched
    assert lib.drw_get_launch_shape(2, 500) == 512


def test_rule_errors(lib):
    assert lib.drw_set_launch_shape(0, 0, 300) == -2
    assert b"256 or 512" in lib.drw_last_error()
    assert lib.drw_set_launch_shape(4, 0, 256) == -2           # no such kernel
    assert lib.drw_set_launch_shape(0, 0, 1) == -2             # 1 (separate kernels) is the fused kernel's only
    assert lib.drw_set_launch_shape(3, 0, 1) == 0
    assert lib.drw_get_launch_shape(-1, 5) == -2
    assert lib.drw_set_launch_shape(1, -1, 256) == 0           # one rule: (0, 256)
    for k in range(7):
        assert lib.drw_set_launch_shape(1, 10 * (k + 1), 512) == 0
    assert lib.drw_set_launch_shape(1, 999, 256) == -2         # a ninth rule
    assert b"rules" in lib.drw_last_error()
    assert lib.drw_set_launch_shape(1, 70, 256) == 0           # replacing one still works


def test_rules_from_timings():
    from deeprwcap_amd.autotune import rules_from_timings
    # measured face-stage times (ms) of the Poisson kernel, 256 / 512 threads
    t = {64: {256: .0269, 512: .0222}, 512: {256: .0312, 512: .0272}, 1024: {256: .0412, 512: .0438},
         2048: {256: .0716, 512: .0747}, 8192: {256: .2434, 512: .2511}}
    assert rules_from_timings(t) == [(0, 512), (1024, 256)]
    # a difference inside the margin does not switch
    assert rules_from_timings({64: {256: 1.0, 512: 0.9}, 128: {256: 0.99, 512: 1.0}}) == [(0, 512)]
    assert rules_from_timings({64: {256: 1.0, 512: 0.9}, 128: {256: 0.9, 512: 1.0}}) == [(0, 512), (128, 256)]


def test_apply_tuning(lib, tmp_path):
    from deeprwcap_amd import autotune
    tuning = {"max_batch": 1024, "launch_shapes": {"face_poisson": [[0, 256], [4096, 512]],
                                                   "face_grad2": [[0, 256]]}}
    p = tmp_path / "tuning.json"
    p.write_text(json.dumps(tuning))
    autotune.apply(str(p))
    assert [lib.drw_get_launch_shape(0, b) for b in (64, 4095, 4096)] == [256, 256, 512]
    assert lib.drw_get_launch_shape(2, 64) == 256
    assert lib.drw_get_launch_shape(1, 64) == 512  # not in the file: built-in rules
