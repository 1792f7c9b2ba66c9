"""Seeded synthetic structure descriptions for the transition query.

A query's dielectric cube is described the way the reference library takes
it (inference_cpp/src/dnn_solver.h:27-30, dnn_solver.cpp:264-272):

* ``layers``  float32 [B, 15, 2]  = (min_z, eps) per planar z-layer, sorted by
  min_z, padded with -9999 (first padded min_z ends the list);
* ``cuboids`` float32 [B, 15, 7]  = (minx, miny, minz, maxx, maxy, maxz, eps),
  first match wins, padded with -9999;

all in normalised cube coordinates [-1, 1].

The generators below produce the workloads used for parity tests and the
benchmark:

* :func:`layered_boxes` - 1-3 planar layers whose normal is drawn from
  {x, y, z} (x/y-normal slabs are expressed as full-span cuboids), interface
  planes uniform on the grid faces of the query's N, eps ~ U[1, 8], plus
  ``0..max_boxes`` axis-aligned box inclusions with eps ~ U[1, 8] whose corners
  lie on grid faces.  All eps are float32 values, so a float64 solve sees
  exactly the same permittivities.
"""
from __future__ import annotations

import numpy as np

PAD = np.float32(-9999.0)
MAX_LAYERS = 15
MAX_CUBOIDS = 15
SPAN = 1.5  # full-span slabs/boxes extend past the cube faces


def empty(B: int):
    layers = np.full((B, MAX_LAYERS, 2), PAD, dtype=np.float32)
    cuboids = np.full((B, MAX_CUBOIDS, 7), PAD, dtype=np.float32)
    return layers, cuboids


def _grid_coord(p, N):
    """Coordinate of grid face p (0..N) of an N-cell cube."""
    return np.float32(-1.0 + 2.0 * p / N)


def layered_boxes(B: int, N, seed: int = 0, max_boxes: int = 2, eps_lo: float = 1.0, eps_hi: float = 8.0,
                  normals=(0, 1, 2)):
    """Generate B structure descriptions.

    ``N`` is an int or an int array [B] (the resolution each query is solved
    at; interfaces are put on that grid).  Returns (layers, cuboids, N[B])."""
    rng = np.random.default_rng(seed)
    Ns = np.broadcast_to(np.asarray(N, dtype=np.int32), (B,)).copy()
    layers, cuboids = empty(B)
    for b in range(B):
        n = int(Ns[b])
        nl = int(rng.integers(1, 4))
        normal = int(normals[int(rng.integers(0, len(normals)))])
        ifaces = np.sort(rng.choice(np.arange(1, n), size=nl - 1, replace=False)) if nl > 1 else np.zeros(0, int)
        eps = rng.uniform(eps_lo, eps_hi, size=nl).astype(np.float32)
        nb = int(rng.integers(0, max_boxes + 1))
        ci = 0
        for _ in range(nb):
            lo = rng.integers(0, n, size=3)
            hi = np.array([rng.integers(lo[d] + 1, n + 1) for d in range(3)])
            e = np.float32(rng.uniform(eps_lo, eps_hi))
            # inclusive box bounds on grid faces: contains cell centres lo..hi-1
            cuboids[b, ci, 0:3] = [_grid_coord(lo[d], n) for d in range(3)]
            cuboids[b, ci, 3:6] = [_grid_coord(hi[d], n) for d in range(3)]
            cuboids[b, ci, 6] = e
            ci += 1
        if normal == 2:
            layers[b, 0] = (-SPAN, eps[0])
            for l in range(1, nl):
                layers[b, l] = (_grid_coord(ifaces[l - 1], n), eps[l])
        else:
            # background = first slab; later slabs as full-span cuboids
            layers[b, 0] = (-SPAN, eps[0])
            for l in range(1, nl):
                lo = _grid_coord(ifaces[l - 1], n)
                hi = _grid_coord(ifaces[l], n) if l < nl - 1 else np.float32(SPAN)
                box = [-SPAN, -SPAN, -SPAN, SPAN, SPAN, SPAN]
                box[normal] = lo
                box[normal + 3] = hi
                cuboids[b, ci, 0:6] = box
                cuboids[b, ci, 6] = eps[l]
                ci += 1
    return layers, cuboids, Ns


def full_descriptions(B: int, N: int, seed: int = 0, eps_lo: float = 1.0, eps_hi: float = 8.0):
    """Descriptions at the format's limits: all 15 z-layers (the first from
    below the cube, the other 14 on distinct grid faces of the N-cell cube,
    sorted) and all 15 cuboids, none padded.  Boxes overlap (first match
    wins), some reach past the cube faces and some are one cell thin.
    Needs N >= 15."""
    if N < MAX_LAYERS:
        raise ValueError("full_descriptions needs N >= 15 (14 distinct interfaces)")
    rng = np.random.default_rng(seed)
    layers, cuboids = empty(B)
    for b in range(B):
        ifaces = np.sort(rng.choice(np.arange(1, N), size=MAX_LAYERS - 1, replace=False))
        eps = rng.uniform(eps_lo, eps_hi, size=MAX_LAYERS).astype(np.float32)
        layers[b, 0] = (-SPAN, eps[0])
        for l in range(1, MAX_LAYERS):
            layers[b, l] = (_grid_coord(ifaces[l - 1], N), eps[l])
        for c in range(MAX_CUBOIDS):
            lo = rng.integers(0, N, size=3)
            hi = np.array([rng.integers(lo[d] + 1, min(N, lo[d] + 1 + (1 if c % 5 == 0 else N)) + 1)
                           for d in range(3)])
            box = [_grid_coord(lo[d], N) for d in range(3)] + [_grid_coord(hi[d], N) for d in range(3)]
            if c % 4 == 1:  # reaches past the cube on one axis
                d = int(rng.integers(0, 3))
                box[d], box[d + 3] = np.float32(-SPAN), np.float32(SPAN)
            cuboids[b, c, 0:6] = box
            cuboids[b, c, 6] = np.float32(rng.uniform(eps_lo, eps_hi))
    return layers, cuboids


def mixed_N(B: int, seed: int = 0, choices=(16, 21, 23)):
    """N drawn uniformly from ``choices`` (deterministic in seed)."""
    rng = np.random.default_rng(seed ^ 0x5EED)
    return rng.choice(np.asarray(choices, dtype=np.int32), size=B)


def wire_crossing(n_wires=8, width=0.1, thick=0.1, length=2.0, pitch=0.2, eps=(3.9, 2.7, 7.0)):
    """Two orthogonal metal layers of `n_wires` parallel wires (µm): layer 1
    along x at z in [0.1, 0.2], layer 2 along y at z in [0.3, 0.4]; planar
    dielectric stack eps[0] below z = 0.25, eps[1] up to z = 0.45, eps[2]
    above; window with 1 µm clearance.  Returns the geometry dict of
    Server.capacitance (conductor ids 0..n-1 layer 1, n..2n-1 layer 2)."""
    span = (n_wires - 1) * pitch + width
    x0 = (length - span) / 2.0
    boxes, ids = [], []
    for i in range(n_wires):
        y = x0 + i * pitch
        boxes.append([0.0, y, 0.1, length, y + width, 0.1 + thick])
        ids.append(i)
    for i in range(n_wires):
        x = x0 + i * pitch
        boxes.append([x, 0.0, 0.3, x + width, length, 0.3 + thick])
        ids.append(n_wires + i)
    window = [-1.0, -1.0, -1.0, length + 1.0, length + 1.0, 1.5]
    return {"window": np.array(window), "layer_z": np.array([-1e9, 0.25, 0.45]), "layer_eps": np.array(eps),
            "boxes": np.array(boxes), "ids": np.array(ids, np.int32)}
