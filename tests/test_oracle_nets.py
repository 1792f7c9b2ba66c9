"""Pins of the FP32 network restatement oracle/nets_ref.py (CPU only).

Two fixtures, both written by tests/golden/make_net_golden.py, hold for each
of the five networks a seeded state dict, an input batch [B,1,1,23,23,23] and
the output of the reference module itself (training_pytorch/src/
architecture.py:11-240, instantiated as the model registry does, main.py:
43-84, in eval mode):
  nets_golden.npz  weights ~ N(0, 1/fan_in), uniform-noise cubes (the
                   seeded-weight set of the GPU parity tests);
  nets_wiring.npz  gain-2 weights, spread batch-norm statistics, layered/box
                   structure cubes: outputs with spatial contrast.
The restatement must reproduce both, which fixes its wiring: the
position-embedding channel order, the depthwise-separable block order and
dilations, the FaceSolverPlus permute, the heads (relu + 1e-10 and sum,
x / sum|x|, softmax, softmax on [:6] with the raw weight in [6]).

The planted-mistake tests change one such detail in the restatement and
check that the wiring fixture then fails by far more than the tolerance, so
the pin is known to detect it.
"""
import os

import numpy as np
import pytest
import torch

from oracle import nets_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-5     # max |restatement - reference module| / max |reference module|
PLANTED = 1e-2  # a planted wiring mistake must move the outputs by at least this


def _load(fname):
    return np.load(os.path.join(ROOT, "tests", "golden", fname))


@pytest.fixture(scope="module")
def gold():
    return _load("nets_wiring.npz")


def _max_err(gold, name):
    """Largest deviation from the recorded reference-module output, relative
    to that output's largest magnitude."""
    P = R.params_from_npz(gold, name)
    x = torch.from_numpy(gold["input"])
    with torch.no_grad():
        y = R.forward(name, x, P).numpy()
    ref = gold[f"{name}/output"]
    assert y.shape == ref.shape, (name, y.shape, ref.shape)
    return float(np.abs(y - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize("fixture", ["nets_golden.npz", "nets_wiring.npz"])
@pytest.mark.parametrize("name", list(R.ARCH))
def test_restatement_matches_reference_modules(fixture, name):
    err = _max_err(_load(fixture), name)
    assert err <= TOL, f"{name} on {fixture}: relative max |nets_ref - reference module| = {err:.2e}"


def test_wiring_fixture_has_contrast(gold):
    """The wiring set's outputs are far from uniform and its relu head clips:
    without that a wrong head or position embedding could go unseen."""
    for name in ("PoissonPredictor", "Gradient1Predictor", "Gradient2Predictor"):
        y = gold[f"{name}/output"].reshape(len(gold["input"]), -1)
        assert ((y.max(1) - y.min(1)) > 0.5 * np.abs(y).max(1)).all(), name
    p = gold["PoissonPredictor/output"].reshape(len(gold["input"]), -1)
    assert ((p < 1e-9).sum(1) > 0).all(), "relu head never clips"
    for name in ("Gradient1Predictor", "Gradient2Predictor"):  # x / sum|x| differs from x / sum x
        assert (gold[f"{name}/output"] < 0).any(), name


def test_golden_detects_swapped_position_channels(gold, monkeypatch):
    real = R.pos_embed

    def swapped(B, H=23, W=23, device="cpu"):
        return real(B, H, W, device).flip(1)

    monkeypatch.setattr(R, "pos_embed", swapped)
    for name in ("PoissonPredictor", "Gradient1Predictor", "Gradient2Predictor"):
        assert _max_err(gold, name) > PLANTED, name


def test_golden_detects_missing_plus_permute(gold, monkeypatch):
    real = R.face_predictor

    def no_permute(x6, P, mode="relu", plus=False, pre="solver."):
        if not plus:
            return real(x6, P, mode, plus, pre)
        x = x6.squeeze(1).squeeze(1)
        y = R.face_solver(x, P, pre, R.FACE_SOLVER_PLUS_BLOCKS).squeeze(1)
        return R.face_head(y, mode)

    monkeypatch.setattr(R, "face_predictor", no_permute)
    assert _max_err(gold, "Gradient2Predictor") > PLANTED


def test_golden_detects_wrong_head(gold, monkeypatch):
    real = R.face_head

    def norm_head_everywhere(x, mode):
        return real(x, "norm")

    monkeypatch.setattr(R, "face_head", norm_head_everywhere)
    assert _max_err(gold, "PoissonPredictor") > PLANTED


def test_golden_detects_reordered_dilations(gold, monkeypatch):
    monkeypatch.setattr(R, "FACE_SOLVER_BLOCKS", [(16, 16, 1), (16, 8, 2), (8, 4, 1), (4, 2, 3)])
    assert _max_err(gold, "PoissonPredictor") > PLANTED
