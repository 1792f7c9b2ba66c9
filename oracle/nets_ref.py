"""FP32 PyTorch restatement of the five transition networks (CPU or GPU).

TEST INFRASTRUCTURE ONLY: the checker for the HIP network kernels (the
"plain PyTorch fp32 reference of the same op"), never part of the product.

Functional forms of training_pytorch/src/architecture.py, written from the
layer tables (eval-mode batch norm, exact erf GELU), taking parameter dicts
keyed like the reference state dicts:
  FaceSolver         architecture.py:44-72   (PoissonPredictor, Gradient1Predictor)
  FaceSolverPlus     architecture.py:75-105  (Gradient2Predictor)
  FacePredictor      architecture.py:108-132 heads 'relu' / 'norm'
  SignedFacePredictor architecture.py:135-173 channel_dim = 1
  FaceSelector       architecture.py:176-205 (PoissonSelector)
  FaceSelectorWeight architecture.py:208-240 (GradientSelectorWeight)
Pinned by tests/test_oracle_nets.py against the outputs of the reference
modules themselves on seeded weights (tests/golden/nets_golden.npz and the
high-contrast tests/golden/nets_wiring.npz, written by
tests/golden/make_net_golden.py): exact match, and each planted wiring slip
(swapped position channels, missing permute, wrong head, reordered
dilations) moves the wiring fixture's outputs by 16 % or more.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

BN_EPS = 1e-5

# (cin, cout, dilation) of the depthwise-separable blocks
FACE_SOLVER_BLOCKS = [(16, 16, 1), (16, 8, 1), (8, 4, 2), (4, 2, 3)]
FACE_SOLVER_PLUS_BLOCKS = [(32, 32, 1), (32, 24, 1), (24, 24, 1), (24, 16, 1), (16, 16, 2), (16, 8, 2), (8, 4, 3)]
ARCH = {
    "PoissonSelector": "selector",
    "PoissonPredictor": "face_relu",
    "GradientSelectorWeight": "selector_weight",
    "Gradient2Predictor": "face_plus_norm",
    "Gradient1Predictor": "face_norm",
}


def _bn(x, P, k):
    return F.batch_norm(x, P[k + ".running_mean"], P[k + ".running_var"], P[k + ".weight"], P[k + ".bias"],
                        training=False, eps=BN_EPS)


def _gelu(x):
    return F.gelu(x)  # exact erf form


def pos_embed(B, H=23, W=23, device="cpu"):
    """Channels [x, y] with values linspace(0, 1) along W and H."""
    y = torch.linspace(0, 1, H, device=device)
    x = torch.linspace(0, 1, W, device=device)
    yy, xx = torch.meshgrid(y, x, indexing="ij")
    return torch.stack((xx, yy), 0).unsqueeze(0).expand(B, 2, H, W)


def face_solver(x, P, pre, blocks):
    """x [B, 23(channels), H, W] -> [B, 1, H, W]."""
    x = torch.cat((x, pos_embed(x.shape[0], x.shape[2], x.shape[3], x.device)), 1)
    x = _gelu(_bn(F.conv2d(x, P[pre + "z_weight.0.weight"]), P, pre + "z_weight.1"))
    for i, (cin, cout, d) in enumerate(blocks):
        k = f"{pre}blocks.{i}.block."
        x = _gelu(_bn(F.conv2d(x, P[k + "0.weight"], padding=d, dilation=d, groups=cin), P, k + "1"))
        x = _gelu(_bn(F.conv2d(x, P[k + "3.weight"]), P, k + "4"))
    return F.conv2d(x, P[pre + "head.weight"], P[pre + "head.bias"])


def face_head(x, mode):
    if mode == "relu":
        x = F.relu(x) + 1e-10
        return x / x.sum(dim=[1, 2], keepdim=True)
    return x / x.abs().sum(dim=[1, 2], keepdim=True)


def face_predictor(x6, P, mode="relu", plus=False, pre="solver."):
    """x6 [B,1,1,23,23,23] (z, y, x) -> [B, 23, 23]."""
    x = x6.squeeze(1).squeeze(1)
    if plus:
        x = x.permute(0, 2, 1, 3)  # channels = y, spatial = (z, x)
    y = face_solver(x, P, pre, FACE_SOLVER_PLUS_BLOCKS if plus else FACE_SOLVER_BLOCKS).squeeze(1)
    return face_head(y, mode)


def selector(x6, P, weight_head=False):
    """FaceSelector / FaceSelectorWeight: [B,1,1,23,23,23] -> [B,6] / [B,7]."""
    x = x6.squeeze(1)
    for i in range(3):
        x = F.conv3d(x, P[f"backbone.{3 * i}.weight"], P[f"backbone.{3 * i}.bias"], stride=2)
        x = _gelu(_bn(x, P, f"backbone.{3 * i + 1}"))
    x = x.flatten(1)
    nlin = 3 if weight_head else 2
    for j in range(nlin):
        x = F.linear(x, P[f"head.{2 * j + 1}.weight"], P[f"head.{2 * j + 1}.bias"])
        if j < nlin - 1:
            x = _gelu(x)
    if weight_head:
        return torch.cat([F.softmax(x[:, :6], dim=1), x[:, 6:]], dim=1)
    return F.softmax(x, dim=1)


def forward(name, x6, P):
    a = ARCH[name]
    if a == "selector":
        return selector(x6, P)
    if a == "selector_weight":
        return selector(x6, P, weight_head=True)
    if a == "face_relu":
        return face_predictor(x6, P, "relu")
    if a == "face_norm":
        return face_predictor(x6, P, "norm")
    return face_predictor(x6, P, "norm", plus=True)


def params_from_npz(d, name, device="cpu"):
    pre = name + "/"
    return {k[len(pre):]: torch.from_numpy(d[k]).to(device) for k in d.files if k.startswith(pre) and
            not k.endswith("/output")}


def query_kernels(kind, layers, cuboids, axis, params, device="cpu"):
    """Full transition kernels of a batch as the reference query computes
    them (dnn_solver_poisson.cpp:86-111, dnn_solver_gradient.cpp:75-164),
    with every face evaluated:

    kind 0 (Poisson): probs [B,6] = PoissonSelector(cube), kern [B,6,529] =
      PoissonPredictor(rotate_poisson_f(cube)).
    kind 1 (gradient): cube_r = axis_rotate(cube); probs, W from
      GradientSelectorWeight(cube_r); kern[f] = post_process(P_f(
      rotate_gradient_f(cube_r)), f, axis) with P_f = Gradient1Predictor for
      f in {0, 1} and Gradient2Predictor otherwise.
    params: name -> state dict (torch tensors).  Returns numpy arrays."""
    import numpy as np
    from oracle import oracle as O
    cubes, _ = O.voxelize(layers, cuboids, 23, normalise=True)
    B = cubes.shape[0]
    kern = np.zeros((B, 6, 529), np.float32)
    with torch.no_grad():
        if kind == 0:
            x = torch.from_numpy(cubes).to(device)[:, None, None]
            probs = forward("PoissonSelector", x, params["PoissonSelector"]).cpu().numpy()
            for f in range(6):
                rot = np.stack([O.rotate_poisson(c, f) for c in cubes])
                y = forward("PoissonPredictor", torch.from_numpy(rot).to(device)[:, None, None],
                            params["PoissonPredictor"]).cpu().numpy()
                kern[:, f] = y.reshape(B, 529)
            return probs, None, kern
        cr = np.stack([O.axis_rotate(cubes[b], int(axis[b])) for b in range(B)])
        out = forward("GradientSelectorWeight", torch.from_numpy(cr).to(device)[:, None, None],
                      params["GradientSelectorWeight"]).cpu().numpy()
        probs, W = out[:, :6], out[:, 6]
        for f in range(6):
            rot = np.stack([O.rotate_gradient(c, f) for c in cr])
            name = "Gradient1Predictor" if f < 2 else "Gradient2Predictor"
            y = forward(name, torch.from_numpy(rot).to(device)[:, None, None], params[name]).cpu().numpy()
            for b in range(B):
                kern[b, f] = O.post_process(y[b], f, int(axis[b])).reshape(-1)
        return probs, W, kern
