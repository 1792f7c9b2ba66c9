"""Held-out error of the PoissonPredictor as the full-kernel error weights it:
with `weighted`, evaluate_face_l1 is the face-probability-weighted per-face
L1, computed here independently from the FP32 restatement of the whole
network query (its own voxelizer and face rotations) on FDM kernels from the
oracle."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_weighted_held_out_is_the_probability_weighted_face_l1(orc):
    import torch
    from deeprwcap_amd import structures as S
    from deeprwcap_amd import training as T
    from deeprwcap_amd import weights as W
    from oracle import nets_ref as R
    n = 6
    L, C, _ = S.layered_boxes(n, 23, seed=321)
    cubes, mx = orc.voxelize(L, C, 23, normalise=True)
    raw, _ = orc.voxelize(L, C, 23, normalise=False)
    G = np.stack([orc.fdm_solve(raw[b].astype(np.float64), orc.GFT, tol=1e-9)[0].reshape(6, 529)
                  for b in range(n)]).astype(np.float32)
    mdir = os.path.join(ROOT, "models")
    params = {m: {k: torch.from_numpy(v) for k, v in W.read_drwnet(os.path.join(mdir, f"{m}.drwnet")).items()}
              for m in W.MODEL_ORDER}
    net = T.FaceNet(plus=False)
    net.load_state_dict(params["PoissonPredictor"], strict=False)
    net.eval()
    idx = np.arange(n)
    got_w = T.evaluate_face_l1("PoissonPredictor", net, cubes.astype(np.float32), G, G, idx, dev="cpu", weighted=True)
    got_u = T.evaluate_face_l1("PoissonPredictor", net, cubes.astype(np.float32), G, G, idx, dev="cpu")
    _, _, rk = R.query_kernels(0, L, C, None, params)
    face = G.sum(-1).astype(np.float64)
    l1 = np.zeros((n, 6))
    for b in range(n):
        for f in range(6):
            l1[b, f] = np.abs(rk[b, f] - G[b, f] / face[b, f]).sum()
    want_w = np.mean([(face[b] / face[b].sum() * l1[b]).sum() for b in range(n)])
    assert abs(got_w - want_w) < 1e-5 * max(1.0, want_w) + 1e-6
    assert abs(got_u - l1.mean()) < 1e-5 + 1e-6
    # the two differ whenever the faces' errors and probabilities differ
    assert abs(want_w - l1.mean()) > 1e-6
