"""Batched device primitives over the C ABI (torch tensors as device memory).

PyTorch only provides device allocations and the stream; all arithmetic runs
in libdnnsolver's HIP kernels.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import check, ptr, stream_handle

GFT, WVTZ, WVTX, WVTY = 0, 1, 2, 3


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.DrwError("no HIP device available")
    return torch


class PackedCubes:
    """Cubes of mixed N packed back to back in one float32 device buffer."""

    def __init__(self, Ns, device="cuda"):
        torch = _torch()
        self.N = np.asarray(Ns, dtype=np.int32)
        n3 = self.N.astype(np.int64) ** 3
        g = 6 * self.N.astype(np.int64) ** 2
        self.eps_off = np.concatenate([[0], np.cumsum(n3)[:-1]]).astype(np.int64)
        self.G_off = np.concatenate([[0], np.cumsum(g)[:-1]]).astype(np.int64)
        self.total_cells = int(n3.sum())
        self.total_panels = int(g.sum())
        self.d_N = torch.from_numpy(self.N).to(device)
        self.d_eps_off = torch.from_numpy(self.eps_off).to(device)
        self.d_G_off = torch.from_numpy(self.G_off).to(device)
        self.eps = torch.empty(self.total_cells, dtype=torch.float32, device=device)
        self.G = torch.empty(self.total_panels, dtype=torch.float64, device=device)

    @property
    def B(self):
        return len(self.N)

    def set_cubes(self, cubes):
        torch = _torch()
        host = np.concatenate([np.asarray(c, np.float32).reshape(-1) for c in cubes])
        self.eps.copy_(torch.from_numpy(host))

    def kernel(self, b):
        n = int(self.N[b])
        o = int(self.G_off[b])
        return self.G[o:o + 6 * n * n]


def fdm_solve(pk: PackedCubes, code, tol=1e-6, order=None, stream=None, want_stats=True):
    """Run the batched FDM kernel on the packed cubes; fills pk.G.

    Returns (iters[B] int32, relres[B] float32) as device tensors."""
    torch = _torch()
    B = pk.B
    d_code = torch.as_tensor(np.broadcast_to(np.asarray(code, np.int32), (B,)).copy()).to(pk.eps.device)
    d_order = None if order is None else torch.as_tensor(np.asarray(order, np.int32)).to(pk.eps.device)
    iters = torch.empty(B, dtype=torch.int32, device=pk.eps.device) if want_stats else None
    relres = torch.empty(B, dtype=torch.float32, device=pk.eps.device) if want_stats else None
    check(_lib.lib().drw_fdm_solve(B, ptr(pk.d_N), ptr(d_code), ptr(pk.eps), ptr(pk.d_eps_off), ptr(pk.G),
                                   ptr(pk.d_G_off), ptr(d_order), float(tol), ptr(iters), ptr(relres),
                                   stream_handle(stream)), "drw_fdm_solve")
    pk._keep = (d_code, d_order)  # keep inputs alive until the stream has consumed them
    return iters, relres


def voxelize(layers, cuboids, N=None, fixed_N=23, normalise=True, out=None, out_off=None, stream=None):
    """Voxelize device (or host) structure arrays; returns (cubes, max)."""
    torch = _torch()
    dev = "cuda"
    L = torch.as_tensor(np.asarray(layers, np.float32) if not hasattr(layers, "data_ptr") else layers).to(dev)
    C = torch.as_tensor(np.asarray(cuboids, np.float32) if not hasattr(cuboids, "data_ptr") else cuboids).to(dev)
    B = L.shape[0]
    d_N = None if N is None else torch.as_tensor(np.asarray(N, np.int32)).to(dev)
    if out is None:
        if N is not None:
            raise ValueError("pass out/out_off for mixed-N voxelization")
        out = torch.empty((B, fixed_N, fixed_N, fixed_N), dtype=torch.float32, device=dev)
    mx = torch.empty(B, dtype=torch.float32, device=dev)
    check(_lib.lib().drw_voxelize(B, ptr(L), ptr(C), ptr(d_N), int(fixed_N), int(bool(normalise)), ptr(out),
                                  ptr(out_off), ptr(mx), stream_handle(stream)), "drw_voxelize")
    return out, mx


def sample_flat(pk: PackedCubes, axis=None, center_eps=None, seed=0, counter=0, stream=None):
    """Flat inverse-CDF next hops from pk.G; returns [B,3] or [B,5] float32."""
    torch = _torch()
    dev = pk.eps.device
    d_axis = None if axis is None else torch.as_tensor(np.asarray(axis, np.int32)).to(dev)
    d_cen = None if center_eps is None else torch.as_tensor(np.asarray(center_eps, np.float32)).to(dev)
    out = torch.empty((pk.B, 3 if axis is None else 5), dtype=torch.float32, device=dev)
    check(_lib.lib().drw_sample_flat(pk.B, ptr(pk.d_N), ptr(d_axis), ptr(pk.G), ptr(pk.d_G_off), ptr(d_cen),
                                     int(seed), int(counter), ptr(out), stream_handle(stream)), "drw_sample_flat")
    return out
