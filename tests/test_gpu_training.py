"""Gradients of the training networks on the GPU (MIOpen convolutions, the
training's own batch norm) against float64 gradients on the CPU for one
batch: the weights pipeline is only as good as these."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _grads(net, x, tgt, dev, dtype):
    import torch
    import torch.nn.functional as F
    net = net.to(dev, dtype).train()
    y = net(x.to(dev, dtype)).reshape(x.shape[0], -1)
    if y.shape[1] == 529:
        p = F.relu(y) + 1e-10
        p = p / p.sum(-1, keepdim=True)
    else:
        p = F.softmax(y[:, :6], -1)
    t = tgt.to(dev, dtype)
    loss = (p - t).abs().sum(-1).mean() + 529.0 * F.mse_loss(p, t)
    net.zero_grad()
    loss.backward()
    return {n: q.grad.detach().double().cpu() for n, q in net.named_parameters()}


@pytest.mark.parametrize("name", ["PoissonPredictor", "Gradient2Predictor", "PoissonSelector"])
def test_net_gradients_match_cpu(name):
    import copy
    import torch
    from deeprwcap_amd import structures as S
    from deeprwcap_amd import training as T
    from deeprwcap_amd import weights as W
    from oracle import oracle as O
    torch.manual_seed(0)
    st = {k: torch.from_numpy(v) for k, v in W.read_drwnet(os.path.join(ROOT, "models", f"{name}.drwnet")).items()}
    net = T.SelectorNet() if name == "PoissonSelector" else T.FaceNet(plus=name == "Gradient2Predictor")
    net.load_state_dict(st, strict=False)
    B = 48
    L, C, _ = S.layered_boxes(B, 23, seed=5)
    cubes, _ = O.voxelize(L, C, 23, normalise=True)
    x = torch.from_numpy(np.stack([O.rotate_poisson(cubes[b], b % 6) for b in range(B)]).astype(np.float32))
    tgt = torch.rand(B, 6 if name == "PoissonSelector" else 529)
    tgt = tgt / tgt.sum(-1, keepdim=True)
    ref = _grads(copy.deepcopy(net).double(), x, tgt, "cpu", torch.float64)
    got = _grads(copy.deepcopy(net), x, tgt, "cuda", torch.float32)
    # per tensor, relative to its own norm or 1e-3 of the whole gradient's
    # (a conv bias ahead of batch norm has an exactly zero gradient)
    total = float(torch.cat([g.reshape(-1) for g in ref.values()]).norm())
    worst = max(float((got[n] - ref[n]).norm()) / max(float(ref[n].norm()), 1e-3 * total) for n in ref)
    print(f"{name}: worst gradient relative error {worst:.2e}", flush=True)
    assert worst < 1e-3


def test_ggft_dataset_files(tmp_path):
    """poisson.bin / gradient.bin written from the GPU FDM (the reference's
    training input format): the records' permittivities are the oracle's
    voxelization of the same seeded descriptions, structure blocks are
    theirs, G^GFT sums to 1 and is non-negative, G^z sums to 0, and two
    records of each file match the oracle's FP64 solve (tolerance 1e-6)."""
    from deeprwcap_amd import structures as S
    from deeprwcap_amd import training as T
    from oracle import oracle as O
    n, seed = 96, 4
    print(f"ggft dataset files: {n} records", flush=True)
    paths = T.ggft_bin_dataset(str(tmp_path), n, seed, chunk=64, log=lambda m: print(m, flush=True))
    L = np.concatenate([S.layered_boxes(b, 23, seed=1_000_003 * (seed + 1) + c0, max_boxes=2)[0]
                        for c0, b in ((0, 64), (64, 32))])
    C = np.concatenate([S.layered_boxes(b, 23, seed=1_000_003 * (seed + 1) + c0, max_boxes=2)[1]
                        for c0, b in ((0, 64), (64, 32))])
    want, _ = O.voxelize(L, C, 23, normalise=False)
    for code in (0, 1):
        cubes, blocks, G = T.read_ggft_bin(paths[code])
        assert len(cubes) == n
        np.testing.assert_array_equal(cubes, want.astype(np.float32).astype(np.float64))
        np.testing.assert_array_equal(blocks, T.description_blocks(L, C))
        s = G.sum(axis=(1, 2))
        if code == 0:
            assert np.abs(s - 1).max() < 1e-4 and G.min() > -1e-7
        else:
            assert np.abs(s).max() < 1e-4
        for b in (0, n - 1):
            ref, _, _ = O.fdm_solve(cubes[b], O.GFT if code == 0 else O.WVTZ, tol=1e-12)
            err = np.abs(G[b].ravel() - ref.ravel()).max() / np.abs(ref).max()
            assert err < 1e-4, (code, b, err)
