"""Weights pipeline: FDM dataset on the GPU, network training, export.

The reference ships its trained networks only as compiled TensorRT engines,
so the weights are produced here, following the reference pipeline
(ggft/src/main.cpp dataset driver -> training_pytorch/src/dataset.py targets
-> training_pytorch/src/train.py losses):

* cubes: seeded structure descriptions (deeprwcap_amd.structures) voxelized
  at N = 23 by the library; kernels G (Poisson, GFT) and G^z (gradient, WVTZ)
  from the library's batched FDM solve;
* targets (dataset.py:108-120): face distribution = per-face sums of G and
  per-face normalised kernels; signed gradient normalised by the face's
  sum |G^z|; face weights [face_tot / sum, sum face_tot];
* augmentation by the query's own face rotations, whose targets follow from
  FDM equivariance: G(rotate_poisson_f(cube))[face 0] = G(cube)[face f], and
  G^z(rotate_gradient_f(cube))[face 1 or 2] = +-G^z(cube)[face f];
* losses: KL for distributions (PoissonSelector, PoissonPredictor), MSE for
  the signed gradients, KL + MSE on W for GradientSelectorWeight
  (main.py:43-84 "mainloss"); AdamW with cosine schedule;
* the network input is eps / max(eps) per cube (train.py:59-60).

Module parameter names equal the reference state-dict keys, so the exported
.drwnet files carry the same tensor names the reference checkpoints have.
"""
from __future__ import annotations

import math
import os
import time

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import device as D
from . import structures as S
from . import weights as W

NET_N = 23


# ------------------------------------------------------------------ modules
class _BatchNorm(nn.modules.batchnorm._BatchNorm):
    """torch.nn.BatchNorm{2,3}d with the same parameters, buffers, state-dict
    names and update rule, computed from plain tensor operations.  On this
    ROCm stack the MIOpen batch-norm kernels that nn.BatchNorm dispatches to
    return gradients that differ from the float64 CPU gradients by up to
    several hundred percent (the face predictors' 4-channel layers; measured,
    tests/test_gpu_training.py); with this form the GPU gradients agree to
    1e-5 while the convolutions stay on MIOpen."""

    def _check_input_dim(self, x):
        if x.dim() not in (4, 5):
            raise ValueError(f"expected 4-D or 5-D input, got {x.dim()}-D")

    def forward(self, x):
        dims = [0] + list(range(2, x.dim()))
        shape = [1, -1] + [1] * (x.dim() - 2)
        if self.training:
            mean = x.mean(dims)
            var = x.var(dims, unbiased=False)
            with torch.no_grad():
                n = x.numel() / x.shape[1]
                self.running_mean.mul_(1 - self.momentum).add_(self.momentum * mean)
                self.running_var.mul_(1 - self.momentum).add_(self.momentum * var * (n / max(n - 1, 1)))
                self.num_batches_tracked.add_(1)
        else:
            mean, var = self.running_mean, self.running_var
        y = (x - mean.view(shape)) * torch.rsqrt(var.view(shape) + self.eps)
        return y * self.weight.view(shape) + self.bias.view(shape)


class _DSB(nn.Module):
    def __init__(self, ci, co, d):
        super().__init__()
        self.block = nn.Sequential(nn.Conv2d(ci, ci, 3, padding=d, dilation=d, groups=ci, bias=False),
                                   _BatchNorm(ci), nn.GELU(), nn.Conv2d(ci, co, 1, bias=False),
                                   _BatchNorm(co), nn.GELU())

    def forward(self, x):
        return self.block(x)


class _Solver(nn.Module):
    def __init__(self, c0, blocks):
        super().__init__()
        self.z_weight = nn.Sequential(nn.Conv2d(25, c0, 1, bias=False), _BatchNorm(c0), nn.GELU())
        self.blocks = nn.Sequential(*[_DSB(ci, co, d) for ci, co, d in blocks])
        self.head = nn.Conv2d(blocks[-1][1], 1, 1, bias=True)
        y = torch.linspace(0, 1, NET_N)
        yy, xx = torch.meshgrid(y, y, indexing="ij")
        self.register_buffer("pe", torch.stack((xx, yy), 0).unsqueeze(0), persistent=False)

    def forward(self, x):
        x = torch.cat((x, self.pe.expand(x.shape[0], 2, NET_N, NET_N)), 1)
        return self.head(self.blocks(self.z_weight(x)))


class FaceNet(nn.Module):
    """FacePredictor / SignedFacePredictor (raw head output; the normalising
    output head is applied by the caller)."""

    def __init__(self, plus=False):
        super().__init__()
        self.plus = plus
        self.solver = _Solver(*(W.FACE_SOLVER_PLUS if plus else W.FACE_SOLVER))

    def forward(self, cube):  # cube [B, 23(z), 23(y), 23(x)]
        x = cube.permute(0, 2, 1, 3) if self.plus else cube
        return self.solver(x).squeeze(1)


class SelectorNet(nn.Module):
    def __init__(self, weighted=False):
        super().__init__()
        ch = [1, 8, 16, 64] if weighted else [1, 4, 8, 16]
        layers = []
        for i in range(3):
            layers += [nn.Conv3d(ch[i], ch[i + 1], 3, stride=2), _BatchNorm(ch[i + 1]), nn.GELU()]
        self.backbone = nn.Sequential(*layers)
        dims = [512, 256, 128, 7] if weighted else [128, 64, 6]
        head = [nn.Flatten()]
        for j in range(len(dims) - 1):
            head.append(nn.Linear(dims[j], dims[j + 1]))
            if j < len(dims) - 2:
                head.append(nn.GELU())
        self.head = nn.Sequential(*head)

    def forward(self, cube):
        return self.head(self.backbone(cube[:, None]))  # logits


def export_state(model: nn.Module) -> dict:
    return {k: v.detach().float().cpu().numpy() for k, v in model.state_dict().items()
            if not k.endswith("num_batches_tracked")}


# ------------------------------------------------------------------ dataset
def fdm_dataset(n: int, seed: int, tol: float = 1e-10, max_boxes: int = 2, chunk: int = 4096, dev="cuda"):
    """Normalised cubes [n,23,23,23] (float16 on the host) and the FDM kernels
    G, G^z [n, 6, 529] (float32), computed with the library on the GPU."""
    cubes = np.empty((n, NET_N, NET_N, NET_N), np.float16)
    G = np.empty((n, 6, 529), np.float32)
    Gz = np.empty((n, 6, 529), np.float32)
    for c0 in range(0, n, chunk):
        b = min(chunk, n - c0)
        # seeds above 10^6: disjoint from the benchmark's and the tests' structures
        L, C, _ = S.layered_boxes(b, NET_N, seed=1_000_003 * (seed + 1) + c0, max_boxes=max_boxes)
        pk = D.PackedCubes([NET_N] * b, device=dev)
        D.voxelize(L, C, fixed_N=NET_N, normalise=True, out=pk.eps.view(b, NET_N, NET_N, NET_N))
        cubes[c0:c0 + b] = pk.eps.view(b, NET_N, NET_N, NET_N).cpu().numpy().astype(np.float16)
        for code, dst in ((0, G), (1, Gz)):
            D.fdm_solve(pk, code, tol=tol, want_stats=False)
            dst[c0:c0 + b] = pk.G.view(b, 6, 529).float().cpu().numpy()
    return cubes, G, Gz


def _rot_table(f, kind):
    """Flat source index of out[z][y][x] for face f's rotation (exact)."""
    n = NET_N
    z, y, x = torch.meshgrid(torch.arange(n), torch.arange(n), torch.arange(n), indexing="ij")
    if kind == "poisson":
        src = {0: (z, y, x), 1: (n - 1 - z, y, x), 2: (y, z, x), 3: (y, n - 1 - z, x), 4: (y, x, z),
               5: (y, x, n - 1 - z)}[f]
    else:
        src = {0: (n - 1 - z, y, x), 1: (z, y, x), 2: (z, y, x), 3: (z, n - 1 - y, x), 4: (z, x, y),
               5: (z, x, n - 1 - y)}[f]
    return (src[0] * n * n + src[1] * n + src[2]).reshape(-1)


def _index_rot(cubes, f, kind):
    """out[z][y][x] = in[src(z,y,x)] for one face f."""
    idx = _rot_table(f, kind).to(cubes.device)
    return cubes.reshape(cubes.shape[0], -1)[:, idx].reshape(cubes.shape)


# ----------------------------------------------------------------- training
def _kl(pred_logp, target):
    t = target.clamp_min(1e-12)
    return (target * (t.log() - pred_logp)).sum(-1).mean()


def train_net(name, cubes, G, Gz, epochs=30, batch=256, lr=1e-3, dev="cuda", log=print, seed=0, minutes=None,
              eval_every=500, n_eval=512, loss_kind="kl", ema=0.999, init=None, best_path=None,
              face_sampling="uniform"):
    """Train one network.  With `minutes`, the run is time-boxed: the number
    of steps is set from the measured step time so that the cosine schedule
    ends when the budget does.  An exponential moving average of the weights
    (decay `ema`, 0 = off) is what gets evaluated and exported.  Every
    `eval_every` steps the held-out error (last 5% of the cubes) is measured
    and the best state is kept.  `init`: a .drwnet to start from (a warm
    restart of the cosine schedule, as CosineAnnealingWarmRestarts does in
    the reference training, train.py:38); `best_path`: a .drwnet rewritten
    with every new best held-out state (a long run keeps its best so far).  `loss_kind` for the PoissonPredictor: "kl"
    (KL + 0.2 L1, the reference's main loss plus an L1 term) or "l1" (L1 +
    MSE of the normalised distribution: bounded gradients, and L1 is the
    quantity the accuracy gate measures); for the PoissonSelector, "kl" (the
    reference's loss) or "l1" (KL + L1 of the face probabilities).  `face_sampling` for the
    PoissonPredictor: "uniform" (every face alike) or "prob" (a cube's face
    is drawn with the probability the selector gives it, the FDM face sum:
    the transition draws the predicted kernel with that probability, so
    this weights the training loss and the held-out error as the
    full-kernel error weights the faces)."""
    torch.manual_seed(seed)
    n = cubes.shape[0]
    ntr = int(n * 0.95)
    Xall = torch.from_numpy(cubes).to(dev)
    Gp = torch.from_numpy(G).to(dev)
    Gg = torch.from_numpy(Gz).to(dev)
    if name in ("PoissonSelector", "GradientSelectorWeight"):
        model = SelectorNet(weighted=name == "GradientSelectorWeight").to(dev)
    else:
        model = FaceNet(plus=name == "Gradient2Predictor").to(dev)
    if init:
        st = {k: torch.from_numpy(v) for k, v in W.read_drwnet(init).items()}
        missing, unexpected = model.load_state_dict(st, strict=False)
        if unexpected or any(not k.endswith("num_batches_tracked") for k in missing):
            raise ValueError(f"{init} does not match {name}: missing {missing}, unexpected {unexpected}")
        model.to(dev)
        log(f"{name}: starting from {init}")
    faces = {"PoissonPredictor": [0, 1, 2, 3, 4, 5], "Gradient1Predictor": [0, 1],
             "Gradient2Predictor": [2, 3, 4, 5]}.get(name, [None])
    opt = torch.optim.AdamW(model.parameters(), lr=lr, weight_decay=1e-10)
    ev_idx = np.arange(ntr, min(n, ntr + n_eval))
    avg = torch.optim.swa_utils.AveragedModel(
        model, multi_avg_fn=torch.optim.swa_utils.get_ema_multi_avg_fn(ema), use_buffers=True) if ema > 0 else None

    # one rotation table per face the network is trained on; every sample of
    # a batch draws its own face, so batch-norm statistics in training are
    # those of the face mix the running statistics describe at inference
    kind_rot = "poisson" if name == "PoissonPredictor" else "gradient"
    rot = torch.stack([_rot_table(f, kind_rot) for f in faces]).to(dev) if faces[0] is not None else None
    face_t = torch.tensor([f if f is not None else -1 for f in faces], device=dev)

    def make_batch(idx, f=None):
        x = Xall[idx].float()
        if name == "PoissonSelector":
            return x, Gp[idx].sum(-1)
        if name == "GradientSelectorWeight":
            tot = Gg[idx].abs().sum(-1)
            s = tot.sum(-1, keepdim=True)
            return x, torch.cat([tot / s, s], 1)
        if face_sampling == "prob" and name == "PoissonPredictor":
            j = torch.multinomial(Gp[idx].sum(-1).clamp_min(0.0), 1).squeeze(1)
        else:
            j = torch.randint(0, len(faces), (idx.shape[0],), device=dev)
        fb = face_t[j]
        xr = torch.gather(x.reshape(x.shape[0], -1), 1, rot[j]).reshape(x.shape)
        if name == "PoissonPredictor":
            k = Gp[idx, fb]
            return xr, k / k.sum(-1, keepdim=True)
        # gradient predictors: target face of the rotated cube; flipping z of
        # an odd-N cube negates the z-gradient kernel (face 0)
        k = Gg[idx, fb] * torch.where(fb == 0, -1.0, 1.0)[:, None]
        return xr, k / k.abs().sum(-1, keepdim=True)

    def loss_of(x, tgt):
        out = model(x)
        if name == "PoissonSelector":
            lp = F.log_softmax(out, -1)
            if loss_kind == "l1":
                return _kl(lp, tgt) + (lp.exp() - tgt).abs().sum(-1).mean()
            return _kl(lp, tgt)
        if name == "GradientSelectorWeight":
            p = F.softmax(out[:, :6], -1)
            return (_kl(F.log_softmax(out[:, :6], -1), tgt[:, :6]) + 0.2 * (p - tgt[:, :6]).abs().sum(-1).mean()
                    + 0.1 * F.mse_loss(out[:, 6], tgt[:, 6]))
        y = out.reshape(out.shape[0], -1)
        if name == "PoissonPredictor":
            p = F.relu(y) + 1e-10
            p = p / p.sum(-1, keepdim=True)
            # the Poisson kernel is positive on every tile: a negative raw output
            # is always wrong and gets no gradient through the relu head, so it
            # is pushed back up directly (keeps the head from dying)
            revive = F.relu(-y).mean() / (y.abs().mean().detach() + 1e-6)
            if loss_kind == "l1":
                return (p - tgt).abs().sum(-1).mean() + 529.0 * F.mse_loss(p, tgt) + 0.1 * revive
            return _kl(p.log(), tgt) + 0.2 * (p - tgt).abs().sum(-1).mean() + 0.1 * revive
        p = y / y.abs().sum(-1, keepdim=True)
        return F.mse_loss(p, tgt) * 529.0 * 50.0 + 0.2 * (p - tgt).abs().sum(-1).mean()

    def held_out():
        m = avg.module if avg is not None else model
        m.eval()
        if name.endswith("Predictor"):
            e = evaluate_face_l1(name, m, cubes, G, Gz, ev_idx, dev=dev, weighted=face_sampling == "prob")
        else:
            with torch.no_grad():
                x = Xall[ev_idx].float()
                out = m(x)
                if name == "PoissonSelector":
                    t = Gp[ev_idx].sum(-1)
                    e = (F.softmax(out, -1) - t / t.sum(-1, keepdim=True)).abs().sum(-1).mean().item()
                else:
                    tot = Gg[ev_idx].abs().sum(-1)
                    e = (F.softmax(out[:, :6], -1) - tot / tot.sum(-1, keepdim=True)).abs().sum(-1).mean().item()
        m.train()
        return e

    def step(st, total):
        idx = torch.randint(0, ntr, (batch,), device=dev)
        x, tgt = make_batch(idx)
        loss = loss_of(x, tgt)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(model.parameters(), 1.0)
        warm = max(1, total // 20)
        frac = st / warm if st < warm else 0.5 * (1 + math.cos(math.pi * (st - warm) / max(1, total - warm)))
        for gr in opt.param_groups:
            gr["lr"] = lr * max(frac, 0.02)
        opt.step()
        if avg is not None:
            avg.update_parameters(model)
        return loss

    best, best_state = float("inf"), None
    if init:
        # a warm restart keeps its starting weights unless training beats them
        # on the held-out cubes
        best = held_out()
        best_state = {k: v.detach().clone() for k, v in model.state_dict().items()}
        log(f"{name}: starting weights held-out L1 {best:.4f}")
    model.train()
    t0 = time.time()
    if minutes is not None:
        # measure the step time after the first steps (library kernel
        # selection and compilation happen there)
        for st in range(30):
            step(st, 1000)
        if dev != "cpu":
            torch.cuda.synchronize()
        t1 = time.time()
        for st in range(30):
            step(st, 1000)
        if dev != "cpu":
            torch.cuda.synchronize()
        per = (time.time() - t1) / 30
        # 25 % margin: the step time measured here underestimates a GPU shared
        # with other trainings
        total = max(100, int((minutes * 60 - (time.time() - t0)) / (1.25 * per * (1 + 2.0 / max(eval_every, 1)))))
        log(f"{name}: {per * 1e3:.1f} ms/step at batch {batch} -> {total} steps")
    else:
        total = epochs * max(1, (ntr * len(faces)) // batch)
    tot_l, nb = 0.0, 0
    for st in range(total):
        loss = step(st, total)
        if st % 50 == 0:
            tot_l += loss.item()
            nb += 1
        if (st + 1) % eval_every == 0 or st == total - 1:
            e = held_out()
            if e < best:
                src = avg.module if avg is not None else model
                best, best_state = e, {k: v.detach().clone() for k, v in src.state_dict().items()}
                if best_path:
                    W.write_drwnet(best_path, {k: v.float().cpu().numpy() for k, v in best_state.items()
                                               if not k.endswith("num_batches_tracked")})
            log(f"{name} step {st + 1}/{total} loss {tot_l / max(nb, 1):.5f} held-out L1 {e:.4f} "
                f"(best {best:.4f}, {time.time() - t0:.0f}s)")
            tot_l, nb = 0.0, 0
    if best_state is not None:
        model.load_state_dict(best_state)
    model.eval()
    return model


def evaluate_face_l1(name, model, cubes, G, Gz, idx, dev="cuda", weighted=False):
    """Mean per-face normalised-kernel L1 on held-out cubes; `weighted`
    (PoissonPredictor): per cube, the faces' L1 weighted by their FDM face
    probabilities, the predictor's share of the full-kernel L1."""
    faces = {"PoissonPredictor": range(6), "Gradient1Predictor": [0, 1], "Gradient2Predictor": [2, 3, 4, 5]}[name]
    Xall = torch.from_numpy(cubes[idx]).to(dev).float()
    errs = []
    with torch.no_grad():
        for f in faces:
            if name == "PoissonPredictor":
                k = torch.from_numpy(G[idx, f]).to(dev)
                tgt = k / k.sum(-1, keepdim=True)
                y = model(_index_rot(Xall, f, "poisson")).reshape(len(idx), -1)
                p = F.relu(y) + 1e-10
                p = p / p.sum(-1, keepdim=True)
            else:
                k = torch.from_numpy(Gz[idx, f]).to(dev)
                k = -k if f == 0 else k
                tgt = k / k.abs().sum(-1, keepdim=True)
                y = model(_index_rot(Xall, f, "gradient")).reshape(len(idx), -1)
                p = y / y.abs().sum(-1, keepdim=True)
            e = (p - tgt).abs().sum(-1)
            if weighted and name == "PoissonPredictor":
                e = e * 6.0 * k.sum(-1) / torch.from_numpy(G[idx]).to(dev).sum((-1, -2))
            errs.append(e.mean().item())
    return float(np.mean(errs))


def full_kernel_l1(sel, pred, x, G, rot):
    """Full-kernel L1 of the Poisson query for cubes x [b,23,23,23]: the
    selector's face probabilities times the predictor's normalised kernel on
    each face's rotated cube, against the FDM kernel G [b,6,529], as
    sum |G_nn - G| / sum |G| per cube (the accuracy gate's measure).  `rot`:
    the six Poisson rotation tables [6, 23^3].  Returns (per-cube L1, G_nn,
    raw predictor output [6b, 529])."""
    b = x.shape[0]
    p = F.softmax(sel(x), -1)                                          # [b, 6]
    flat = x.reshape(b, -1)
    xr = torch.stack([flat[:, rot[f]] for f in range(6)], 1).reshape(6 * b, *x.shape[1:])
    y = pred(xr).reshape(b, 6, -1)
    k = F.relu(y) + 1e-10
    k = k / k.sum(-1, keepdim=True)
    Gn = p[:, :, None] * k
    return (Gn - G).abs().sum((1, 2)) / G.abs().sum((1, 2)), Gn, y


def train_joint_poisson(cubes, G, sel_init, pred_init, minutes=10.0, batch=64, lr_sel=2e-4, lr_pred=2e-4,
                        dev="cuda", log=print, seed=0, eval_every=250, n_eval=2048, ema=0.999, out_dir=None):
    """Fine-tune the PoissonSelector and PoissonPredictor together on the
    quantity the accuracy gate measures: per cube, the full 6x529 kernel
    p_f * k_f of the two networks against the normalised FDM kernel, with
    loss L1 + 3174 * MSE (plus the predictor's relu-head revival term).  Each
    step runs the selector on `batch` cubes and the predictor on all six
    rotated views of each (batch-norm statistics of the six-face mix the
    running statistics describe at inference).  Both networks start from the
    given .drwnet files; the held-out full-kernel L1 of the starting pair is
    measured first and a pair is kept only when it beats it.  With
    `out_dir`, every new best pair is written there as
    PoissonSelector.drwnet and PoissonPredictor.drwnet.  Returns (best
    held-out L1, starting held-out L1)."""
    torch.manual_seed(seed)
    n = cubes.shape[0]
    ntr = int(n * 0.95)
    Xall = torch.from_numpy(cubes).to(dev)
    Gp = torch.from_numpy(G).to(dev)
    sel, pred = SelectorNet().to(dev), FaceNet().to(dev)
    for m, path in ((sel, sel_init), (pred, pred_init)):
        st = {k: torch.from_numpy(v) for k, v in W.read_drwnet(path).items()}
        missing, unexpected = m.load_state_dict(st, strict=False)
        if unexpected or any(not k.endswith("num_batches_tracked") for k in missing):
            raise ValueError(f"{path}: missing {missing}, unexpected {unexpected}")
    rot = torch.stack([_rot_table(f, "poisson") for f in range(6)]).to(dev)
    opt = torch.optim.AdamW([{"params": sel.parameters(), "lr": lr_sel, "base": lr_sel},
                             {"params": pred.parameters(), "lr": lr_pred, "base": lr_pred}], weight_decay=1e-10)
    avgs = [torch.optim.swa_utils.AveragedModel(m, multi_avg_fn=torch.optim.swa_utils.get_ema_multi_avg_fn(ema),
                                                use_buffers=True) if ema > 0 else None for m in (sel, pred)]
    ev = torch.arange(ntr, min(n, ntr + n_eval), device=dev)

    def current():
        return [a.module if a is not None else m for a, m in zip(avgs, (sel, pred))]

    def held_out():
        s, q = current()
        s.eval()
        q.eval()
        errs = []
        with torch.no_grad():
            for i in range(0, ev.shape[0], 256):
                j = ev[i:i + 256]
                errs.append(full_kernel_l1(s, q, Xall[j].float(), Gp[j], rot)[0])
        s.train()
        q.train()
        return torch.cat(errs).mean().item()

    def save(states):
        if out_dir:
            os.makedirs(out_dir, exist_ok=True)
            for name, st in zip(("PoissonSelector", "PoissonPredictor"), states):
                W.write_drwnet(os.path.join(out_dir, f"{name}.drwnet"),
                               {k: v.float().cpu().numpy() for k, v in st.items()
                                if not k.endswith("num_batches_tracked")})

    def step(st, total):
        idx = torch.randint(0, ntr, (batch,), device=dev)
        g = Gp[idx]
        g = g / g.sum((1, 2), keepdim=True)
        l1, Gn, y = full_kernel_l1(sel, pred, Xall[idx].float(), g, rot)
        revive = F.relu(-y).mean() / (y.abs().mean().detach() + 1e-6)
        loss = l1.mean() + 3174.0 * F.mse_loss(Gn, g) + 0.1 * revive
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(list(sel.parameters()) + list(pred.parameters()), 1.0)
        warm = max(1, total // 20)
        frac = st / warm if st < warm else 0.5 * (1 + math.cos(math.pi * (st - warm) / max(1, total - warm)))
        for gr in opt.param_groups:
            gr["lr"] = gr["base"] * max(frac, 0.02)
        opt.step()
        for a, m in zip(avgs, (sel, pred)):
            if a is not None:
                a.update_parameters(m)
        return loss

    start = held_out()
    best = start
    best_states = [{k: v.detach().clone() for k, v in m.state_dict().items()} for m in (sel, pred)]
    log(f"joint Poisson: starting pair held-out full-kernel L1 {start:.5f} ({ev.shape[0]} cubes)")
    t0 = time.time()
    for st in range(20):
        step(st, 1000)
    if dev != "cpu":
        torch.cuda.synchronize()
    t1 = time.time()
    for st in range(20):
        step(st, 1000)
    if dev != "cpu":
        torch.cuda.synchronize()
    per = (time.time() - t1) / 20
    total = max(50, int((minutes * 60 - (time.time() - t0)) / (1.2 * per)))
    log(f"joint Poisson: {per * 1e3:.1f} ms/step at batch {batch} -> {total} steps")
    tot_l, nb = 0.0, 0
    for st in range(total):
        loss = step(st, total)
        if st % 50 == 0:
            tot_l += loss.item()
            nb += 1
        if (st + 1) % eval_every == 0 or st == total - 1:
            e = held_out()
            if e < best:
                best = e
                best_states = [{k: v.detach().clone() for k, v in m.state_dict().items()} for m in current()]
                save(best_states)
            log(f"joint step {st + 1}/{total} loss {tot_l / max(nb, 1):.5f} held-out {e:.5f} (best {best:.5f}, "
                f"start {start:.5f}, {time.time() - t0:.0f}s)")
            tot_l, nb = 0.0, 0
    save(best_states)
    return best, start


# --------------------------------------------------------------------------
# GGFT dataset files (poisson.bin / gradient.bin, the reference's training
# input): a float64 header [N, block_width], then per sample the N^3
# permittivities (x fastest), GGFT_STRUCTURES blocks of 7 structure values and
# the 6 N^2 kernel values, all float64 (ggft/src/ggft_solver.cpp:35-48,
# ggft_solver.h:180-191, ggft/src/main.cpp:59-78; read by
# training_pytorch/src/dataset.py:55-86 with n_structures = 16 and by
# DNNSolver::readBinaryFile with structureSize = 112).

GGFT_STRUCTURES = 16  # 15 blocks + the background (dielectric_generator.h:55-79)


def description_blocks(layers, cuboids, n=GGFT_STRUCTURES):
    """Structure blocks [B, n, 7] of descriptions in the generator's record
    form (centre x, y, z, half-extent x, y, z, value; dielectric_generator.h:
    8-19, 95-106), first matching block wins as there: the cuboids in listing
    order, then each layer as a slab over the whole cube (x, y half-extent 2,
    the generator's span; the first layer reaches down to -2), zero blocks
    after."""
    L = np.asarray(layers, np.float64).reshape(len(layers), -1, 2)
    C = np.asarray(cuboids, np.float64).reshape(len(cuboids), -1, 7)
    out = np.zeros((len(L), n, 7))
    for b in range(len(L)):
        rows = []
        for c in C[b]:
            if c[0] == S.PAD:
                break
            lo, hi = np.maximum(c[:3], -2.0), np.minimum(c[3:6], 2.0)
            rows.append([*(lo + hi) / 2, *(hi - lo) / 2, c[6]])
        lay = [l for l in L[b] if l[0] != S.PAD]
        for k, (z0, e) in enumerate(lay):
            lo = -2.0 if k == 0 else max(z0, -2.0)  # the first layer holds everything below the next
            hi = min(lay[k + 1][0], 2.0) if k + 1 < len(lay) else 2.0
            rows.append([0.0, 0.0, (lo + hi) / 2, 2.0, 2.0, (hi - lo) / 2, e])
        if len(rows) > n:
            raise ValueError(f"description {b} needs {len(rows)} structure blocks, a record holds {n}")
        out[b, :len(rows)] = rows
    return out


def write_ggft_bin(path, cubes, G, blocks=None, N=NET_N, block_width=1, append=False):
    """Write (or append, without a second header) GGFT dataset records:
    cubes [n, N, N, N] (x fastest), blocks [n, GGFT_STRUCTURES, 7] (zeros when
    None), G [n, 6, N*N], all stored as float64."""
    cubes = np.asarray(cubes).reshape(-1, N ** 3)
    G = np.asarray(G).reshape(len(cubes), 6 * N * N)
    if blocks is None:
        blocks = np.zeros((len(cubes), GGFT_STRUCTURES, 7))
    blocks = np.asarray(blocks, np.float64).reshape(len(cubes), -1)
    rec = np.concatenate([cubes.astype(np.float64), blocks, G.astype(np.float64)], axis=1)
    with open(path, "ab" if append else "wb") as f:
        if not append:
            f.write(np.array([N, block_width], np.float64).tobytes())
        f.write(np.ascontiguousarray(rec).tobytes())


def read_ggft_bin(path, N=NET_N, n_structures=GGFT_STRUCTURES):
    """(cubes [n, N, N, N], blocks [n, n_structures, 7], G [n, 6, N*N]),
    float64, memory-mapped, from a GGFT dataset file (dataset.py:55-86)."""
    v = np.memmap(path, dtype=np.float64, mode="r")
    n, bw = int(v[0]), int(v[1])
    if n != N or N % bw:
        raise ValueError(f"{path}: header N={n}, block width {bw}; expected N={N}")
    bn = N // bw
    rec = bn ** 3 + 7 * n_structures + 6 * N * N
    if (v.size - 2) % rec:
        raise ValueError(f"{path}: {v.size - 2} values are no whole number of {rec}-value records")
    body = v[2:].reshape(-1, rec)
    return (body[:, :bn ** 3].reshape(-1, bn, bn, bn), body[:, bn ** 3:bn ** 3 + 7 * n_structures]
            .reshape(-1, n_structures, 7), body[:, -6 * N * N:].reshape(-1, 6, N * N))


def ggft_bin_dataset(out_dir, n, seed, tol=1e-6, max_boxes=2, chunk=4096, dev="cuda", log=print):
    """poisson.bin (G^GFT) and gradient.bin (G^z, code WVTZ) of n seeded
    descriptions under out_dir, as the reference's GGFT driver writes them
    (ggft/run_ggft.sh, main.cpp:38-78: solver tolerance 1e-6), solved by the
    library's FDM on the GPU; DIEL holds the unnormalised permittivities."""
    os.makedirs(out_dir, exist_ok=True)
    paths = {0: os.path.join(out_dir, "poisson.bin"), 1: os.path.join(out_dir, "gradient.bin")}
    for c0 in range(0, n, chunk):
        b = min(chunk, n - c0)
        L, C, _ = S.layered_boxes(b, NET_N, seed=1_000_003 * (seed + 1) + c0, max_boxes=max_boxes)
        blocks = description_blocks(L, C)
        pk = D.PackedCubes([NET_N] * b, device=dev)
        D.voxelize(L, C, fixed_N=NET_N, normalise=False, out=pk.eps.view(b, NET_N, NET_N, NET_N))
        cubes = pk.eps.view(b, NET_N, NET_N, NET_N).cpu().numpy()
        for code in (0, 1):
            D.fdm_solve(pk, code, tol=tol, want_stats=False)
            write_ggft_bin(paths[code], cubes, pk.G.view(b, 6, 529).cpu().numpy(), blocks, append=c0 > 0)
        log(f"ggft records: {c0 + b}/{n}")
    return paths


def save_dataset(path, cubes, G, Gz):
    np.save(path + ".cubes.npy", cubes)
    np.save(path + ".G.npy", G)
    np.save(path + ".Gz.npy", Gz)


def load_dataset(path):
    return (np.load(path + ".cubes.npy", mmap_mode="r"), np.load(path + ".G.npy", mmap_mode="r"),
            np.load(path + ".Gz.npy", mmap_mode="r"))


SHARE = {"PoissonSelector": 0.12, "PoissonPredictor": 0.38, "GradientSelectorWeight": 0.10,
         "Gradient2Predictor": 0.24, "Gradient1Predictor": 0.16}


def train_one(out_dir, name, cubes, G, Gz, minutes, batch, lr, seed, log, loss_kind="kl", ema=0.999, init=None,
              face_sampling="uniform"):
    os.makedirs(out_dir, exist_ok=True)
    m = train_net(name, np.ascontiguousarray(cubes), np.ascontiguousarray(G), np.ascontiguousarray(Gz), dev="cuda",
                  log=log, batch=batch, lr=lr, seed=seed, minutes=minutes, loss_kind=loss_kind, ema=ema, init=init,
                  best_path=os.path.join(out_dir, f"{name}.drwnet"), face_sampling=face_sampling)
    W.write_drwnet(os.path.join(out_dir, f"{name}.drwnet"), export_state(m))
    return m


def train_all(out_dir, n_cubes=40000, epochs=30, seed=0, log=print, dev="cuda", minutes=None, batch=256, lr=1e-3):
    """Sequential training of the five networks (`minutes` split by SHARE)."""
    t0 = time.time()
    cubes, G, Gz = fdm_dataset(n_cubes, seed, dev=dev)
    log(f"dataset: {n_cubes} cubes, FDM kernels in {time.time() - t0:.1f}s")
    states = {}
    for name in W.MODEL_ORDER:
        m = train_net(name, cubes, G, Gz, epochs=epochs, dev=dev, log=log, batch=batch, lr=lr,
                      minutes=None if minutes is None else minutes * SHARE[name])
        states[name] = export_state(m)
        os.makedirs(out_dir, exist_ok=True)
        W.write_drwnet(os.path.join(out_dir, f"{name}.drwnet"), states[name])  # checkpoint per network
    path = W.export_model_dir(out_dir, states)
    log(f"exported {path} in {time.time() - t0:.0f}s")
    return path


def train_parallel(out_dir, data, minutes, batch, lr, seed, log=print):
    """One process per network on the same GPU, each for the whole budget:
    the networks are small, so a single training step leaves the GPU mostly
    idle and five concurrent processes train about five times as much."""
    import subprocess
    import sys
    procs = []
    for i, name in enumerate(W.MODEL_ORDER):
        cmd = [sys.executable, "-u", "-m", "deeprwcap_amd.training", "--out", out_dir, "--data", data, "--only", name,
               "--minutes", str(minutes), "--batch", str(batch), "--lr", str(lr), "--seed", str(seed + i)]
        procs.append((name, subprocess.Popen(cmd)))
    rc = 0
    for name, p in procs:
        r = p.wait()
        log(f"{name}: exit {r}")
        rc = rc or r
    if rc:
        raise RuntimeError("a training process failed")
    states = {n: W.read_drwnet(os.path.join(out_dir, f"{n}.drwnet")) for n in W.MODEL_ORDER}
    return W.export_model_dir(out_dir, states)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="FDM dataset + training + .drwnet export")
    ap.add_argument("--out", help="model directory (models.txt + .drwnet files)")
    ap.add_argument("--n", type=int, default=40000, help="cubes in the FDM dataset")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--minutes", type=float, default=None,
                    help="wall-clock budget (sequential: for all five nets; --parallel/--only: per net)")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--loss", default="kl", choices=("kl", "l1"), help="PoissonPredictor / PoissonSelector loss")
    ap.add_argument("--ema", type=float, default=0.999, help="weight EMA decay (0: off)")
    ap.add_argument("--face-sampling", default="uniform", choices=("uniform", "prob"),
                    help="PoissonPredictor: faces drawn uniformly or with their FDM face probability")
    ap.add_argument("--make-data", metavar="PATH", help="only build the FDM dataset and save it under PATH.*.npy")
    ap.add_argument("--make-bin", metavar="DIR", help="only write DIR/poisson.bin and DIR/gradient.bin in the GGFT "
                    "dataset format (the reference's training input)")
    ap.add_argument("--data", metavar="PATH", help="use a dataset saved by --make-data")
    ap.add_argument("--only", help="train this network only (needs --data)")
    ap.add_argument("--init", metavar="DRWNET", help="with --only: start from these weights")
    ap.add_argument("--parallel", action="store_true", help="one process per network (needs --data)")
    ap.add_argument("--joint", nargs=2, metavar=("SELECTOR", "PREDICTOR"),
                    help="fine-tune a PoissonSelector and PoissonPredictor (.drwnet) together on the full-kernel "
                         "L1 (needs --data; --lr for both, --batch cubes per step)")
    args = ap.parse_args()

    def _log(msg):
        print(msg, flush=True)

    if args.make_bin:
        t0 = time.time()
        paths = ggft_bin_dataset(args.make_bin, args.n, args.seed, log=_log)
        _log(f"GGFT dataset: {args.n} records in {time.time() - t0:.1f}s -> {paths[0]}, {paths[1]}")
    elif args.make_data:
        t0 = time.time()
        save_dataset(args.make_data, *fdm_dataset(args.n, args.seed))
        _log(f"dataset: {args.n} cubes in {time.time() - t0:.1f}s -> {args.make_data}.*.npy")
    elif args.joint:
        cubes, G, _ = load_dataset(args.data)
        best, start = train_joint_poisson(cubes, G, *args.joint, minutes=args.minutes or 10.0, batch=args.batch,
                                          lr_sel=args.lr, lr_pred=args.lr, log=_log, seed=args.seed, ema=args.ema,
                                          out_dir=args.out)
        _log(f"joint Poisson: held-out full-kernel L1 {start:.5f} -> {best:.5f}")
    elif args.only:
        train_one(args.out, args.only, *load_dataset(args.data), args.minutes, args.batch, args.lr, args.seed, _log,
                  loss_kind=args.loss, ema=args.ema, init=args.init, face_sampling=args.face_sampling)
    elif args.parallel:
        _log(f"exported {train_parallel(args.out, args.data, args.minutes, args.batch, args.lr, args.seed, _log)}")
    else:
        train_all(args.out, n_cubes=args.n, epochs=args.epochs, seed=args.seed, log=_log, minutes=args.minutes,
                  batch=args.batch, lr=args.lr)
