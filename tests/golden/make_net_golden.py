"""Golden vectors for the five networks' wiring.

Imports the reference architecture module (training_pytorch/src/architecture.py,
which needs only torch) in this container, instantiates each network the way
the model registry does (training_pytorch/src/main.py:43-84), gives it seeded
random weights and batch-norm statistics, and records, in eval mode, the
state dict (inputs), a seeded input batch and the forward output (expected
output).  Only these numbers are kept; no reference source is copied.

    python tests/golden/make_net_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_SRC = "/root/reference/training_pytorch/src"
OUT = os.path.join(ROOT, "tests", "golden", "nets_golden.npz")


def randomise(model, gen):
    with torch.no_grad():
        for name, p in model.named_parameters():
            fan = p[0].numel() if p.dim() > 1 else 1
            p.copy_(torch.randn(p.shape, generator=gen) / np.sqrt(max(fan, 1)))
            if name.endswith(".weight") and p.dim() == 1:  # batch-norm gamma
                p.copy_(0.5 + torch.rand(p.shape, generator=gen))
        for name, b in model.named_buffers():
            if name.endswith("running_mean"):
                b.copy_(0.2 * torch.randn(b.shape, generator=gen))
            elif name.endswith("running_var"):
                b.copy_(0.5 + torch.rand(b.shape, generator=gen))
    model.eval()
    return model


def main():
    sys.path.insert(0, REF_SRC)
    import architecture as A  # noqa: E402  (reference module, imported here only)
    nets = {
        "PoissonSelector": lambda: A.FaceSelector(),
        "PoissonPredictor": lambda: A.FacePredictor(head_mode="relu"),
        "GradientSelectorWeight": lambda: A.FaceSelectorWeight(),
        "Gradient2Predictor": lambda: A.SignedFacePredictor(head_mode="norm"),
        "Gradient1Predictor": lambda: A.FacePredictor(head_mode="norm"),
    }
    gen = torch.Generator().manual_seed(1234)
    x = torch.rand((6, 1, 1, 23, 23, 23), generator=gen)
    x = x / x.amax(dim=(-3, -2, -1), keepdim=True)
    arrays = {"input": x.numpy()}
    for name, mk in nets.items():
        m = randomise(mk(), gen)
        with torch.no_grad():
            y = m(x)
        arrays[f"{name}/output"] = y.numpy()
        for k, v in m.state_dict().items():
            if k.endswith("num_batches_tracked"):
                continue
            arrays[f"{name}/{k}"] = v.numpy()
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, "entries", len(arrays))


if __name__ == "__main__":
    main()
