""".drwnet weight files and the shipped model list: the format round-trips
bit-exactly, models/models.txt lists the five networks in the reference's
index order (inference_cpp/models.txt:4-8), and every shipped file holds
exactly the state dict of its architecture (training_pytorch/src/
architecture.py:44-240), whose learnable parameter counts are those of the
reference networks: 13,158 / 1,403 / 196,711 / 5,957 / 1,403."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PARAMS = {"PoissonSelector": 13158, "PoissonPredictor": 1403, "GradientSelectorWeight": 196711,
          "Gradient2Predictor": 5957, "Gradient1Predictor": 1403}


def _net(name):
    from deeprwcap_amd import training as T
    if name in ("PoissonSelector", "GradientSelectorWeight"):
        return T.SelectorNet(weighted=name == "GradientSelectorWeight")
    return T.FaceNet(plus=name == "Gradient2Predictor")


@pytest.mark.parametrize("name", list(PARAMS))
def test_drwnet_round_trip(tmp_path, name):
    from deeprwcap_amd import weights as W
    st = W.random_state(name, seed=3)
    p = str(tmp_path / f"{name}.drwnet")
    W.write_drwnet(p, st)
    back = W.read_drwnet(p)
    assert list(back) == list(st)
    for k in st:
        assert back[k].dtype == np.float32 and back[k].shape == np.asarray(st[k]).shape
        np.testing.assert_array_equal(back[k], np.asarray(st[k], np.float32))


def test_model_list_order():
    from deeprwcap_amd import weights as W
    lines = [l.split("#")[0].strip() for l in open(os.path.join(ROOT, "models", "models.txt"))]
    paths = [l for l in lines if l]
    assert paths == [f"{n}.drwnet" for n in W.MODEL_ORDER]
    assert W.MODEL_ORDER == ["PoissonSelector", "PoissonPredictor", "GradientSelectorWeight", "Gradient2Predictor",
                             "Gradient1Predictor"]


@pytest.mark.parametrize("name", list(PARAMS))
def test_shipped_weights_match_architecture(name):
    import torch
    from deeprwcap_amd import weights as W
    st = W.read_drwnet(os.path.join(ROOT, "models", f"{name}.drwnet"))
    net = _net(name)
    ref = {k: v for k, v in net.state_dict().items() if not k.endswith("num_batches_tracked")}
    assert set(st) == set(ref)
    for k, v in ref.items():
        assert st[k].shape == tuple(v.shape), k
        assert np.isfinite(st[k]).all(), k
    assert sum(p.numel() for p in net.parameters()) == PARAMS[name]
    net.load_state_dict({k: torch.from_numpy(v) for k, v in st.items()}, strict=False)
