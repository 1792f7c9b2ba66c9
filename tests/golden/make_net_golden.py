"""Golden vectors for the five networks' wiring.

Imports the reference architecture module (training_pytorch/src/architecture.py,
which needs only torch) in this container, instantiates each network the way
the model registry does (training_pytorch/src/main.py:43-84), gives it seeded
random weights and batch-norm statistics, and records, in eval mode, the
state dict (inputs), a seeded input batch and the forward output (expected
output).  Only these numbers are kept; no reference source is copied.

    python tests/golden/make_net_golden.py            # nets_golden.npz
    python tests/golden/make_net_golden.py contrast   # nets_wiring.npz

nets_golden.npz (weights ~ N(0, 1/fan_in), uniform-noise cubes) is the
seeded-weight set of the GPU parity tests; its outputs are close to uniform,
so it cannot tell some wiring slips apart.  nets_wiring.npz pins the wiring:
larger weights (gain 2), spread batch-norm statistics and layered/box
structure cubes give outputs with real spatial contrast and negative raw head
values, so a swapped position-embedding channel, a missing permute, a wrong
head or reordered dilations change them by far more than rounding.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_SRC = "/root/reference/training_pytorch/src"
OUT = os.path.join(ROOT, "tests", "golden", "nets_golden.npz")
OUT_WIRING = os.path.join(ROOT, "tests", "golden", "nets_wiring.npz")


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


def randomise_contrast(model, gen):
    with torch.no_grad():
        for name, p in model.named_parameters():
            fan = p[0].numel() if p.dim() > 1 else 1
            if name.endswith(".weight") and p.dim() == 1:  # batch-norm gamma
                p.copy_(0.5 + 1.5 * torch.rand(p.shape, generator=gen))
            elif p.dim() == 1:  # biases, batch-norm beta
                p.copy_(0.5 * torch.randn(p.shape, generator=gen))
            else:
                p.copy_(2.0 * torch.randn(p.shape, generator=gen) / np.sqrt(max(fan, 1)))
            # bf16-representable values: the float32 fixture then compresses well
            p.copy_(p.to(torch.bfloat16).float())
        for name, b in model.named_buffers():
            if name.endswith("running_mean"):
                b.copy_(0.1 * torch.randn(b.shape, generator=gen))
            elif name.endswith("running_var"):
                b.copy_(0.05 + 0.3 * torch.rand(b.shape, generator=gen))
            if name.endswith(("running_mean", "running_var")):
                b.copy_(b.to(torch.bfloat16).float())
    model.eval()
    return model


def structure_cubes(B, seed):
    """Normalised permittivity cubes [B,1,1,23,23,23] of seeded layered/box
    structures (the description format the query path voxelizes)."""
    sys.path.insert(0, ROOT)
    from deeprwcap_amd import structures
    from oracle import oracle as O
    L, C, _ = structures.layered_boxes(B, 23, seed=seed, max_boxes=2)
    cubes, _ = O.voxelize(L, C, 23, normalise=True)
    return torch.from_numpy(np.ascontiguousarray(cubes, np.float32)).reshape(B, 1, 1, 23, 23, 23)


def main(mode=""):
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
    out, rnd = OUT, randomise
    if mode == "contrast":
        gen = torch.Generator().manual_seed(4321)
        x = structure_cubes(5, seed=99)
        out, rnd = OUT_WIRING, randomise_contrast
    arrays = {"input": x.numpy()}
    for name, mk in nets.items():
        m = rnd(mk(), gen)
        with torch.no_grad():
            y = m(x)
        arrays[f"{name}/output"] = y.numpy()
        for k, v in m.state_dict().items():
            if k.endswith("num_batches_tracked"):
                continue
            arrays[f"{name}/{k}"] = v.numpy()
    np.savez_compressed(out, **arrays)
    print("wrote", out, "entries", len(arrays))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "")
