"""Reader for the cap3d structure format of the reference's test cases
(testcases/cap3d/*.cap3d; semantics of the fields as the reference's
executable/run_script.py:32-76 drives them) and the end-to-end capacitance
harness around the walker.

Recognised sections:

* ``<window>`` v1(...) v2(...): the simulation box (potential 0 outside);
* ``<plate_medium>`` diel, z_top: planar dielectric layers, each from the
  previous z_top (or -inf) up to its own z_top;
* ``<medium>`` with ``<block>``s and diel: dielectric boxes (they override the
  planar layers where they lie);
* ``<conductor>`` name with ``<block>``s: conductor boxes;
* ``<task>`` ``<capacitance>`` names: the master nets;
* ``<layer>``, ``<parameter>``: read and ignored.

A block is basepoint(x,y,z) + v1 + v2 + hvector; the test cases only use
axis-aligned edge vectors, which is required here (other blocks raise).
Coordinates are in micrometres; capacitances come out in farads.
"""
from __future__ import annotations

import re

import numpy as np

EPS0_F_PER_UM = 8.8541878128e-18  # vacuum permittivity, F/um

_VEC = re.compile(r"(basepoint|v1|v2|hvector)\s*\(\s*([^)]*)\)")


def _box(block_text: str):
    vals = {m.group(1): np.array([float(t) for t in m.group(2).split(",")]) for m in _VEC.finditer(block_text)}
    for k in ("basepoint", "v1", "v2", "hvector"):
        if k not in vals or vals[k].size != 3:
            raise ValueError(f"cap3d block without {k}: {block_text.strip()[:120]}")
    edges = np.stack([vals["v1"], vals["v2"], vals["hvector"]])
    if ((edges != 0).sum(1) != 1).any() or ((edges != 0).sum(0) != 1).any():
        raise ValueError(f"cap3d block is not axis-aligned: {block_text.strip()[:120]}")
    p0 = vals["basepoint"]
    corners = [p0, p0 + edges.sum(0)]
    lo, hi = np.minimum(*corners), np.maximum(*corners)
    return [*lo, *hi]


def _sections(text: str, tag: str):
    return re.findall(rf"<{tag}>(.*?)</{tag}>", text, flags=re.S)


def _field(body: str, key: str):
    m = re.search(rf"\b{key}\s+([^\s<]+)", body)
    return m.group(1) if m else None


def parse(text: str) -> dict:
    """Structure of a cap3d file: window, planar layers, dielectric boxes,
    conductors (name -> boxes) and master net names."""
    win = _sections(text, "window")
    if len(win) != 1:
        raise ValueError("cap3d: expected one <window>")
    w = {m.group(1): [float(t) for t in m.group(2).split(",")] for m in _VEC.finditer(win[0])}
    window = [*np.minimum(w["v1"], w["v2"]), *np.maximum(w["v1"], w["v2"])]
    plates = []
    for body in _sections(text, "plate_medium"):
        plates.append((float(_field(body, "z_top")), float(_field(body, "diel"))))
    plates.sort()
    # layer i covers (z_top[i-1], z_top[i]]; the first starts at -inf
    layer_z = [-1e30] + [z for z, _ in plates[:-1]]
    layer_eps = [e for _, e in plates]
    media_box, media_eps = [], []
    for body in _sections(text, "medium"):
        eps = float(re.findall(r"\bdiel\s+([^\s<]+)", re.sub(r"<block>.*?</block>", "", body, flags=re.S))[-1])
        for blk in _sections(body, "block"):
            media_box.append(_box(blk))
            media_eps.append(eps)
    names, boxes, ids = [], [], []
    for body in _sections(text, "conductor"):
        name = _field(re.sub(r"<block>.*?</block>", "", body, flags=re.S), "name")
        names.append(name)
        for blk in _sections(body, "block"):
            boxes.append(_box(blk))
            ids.append(len(names) - 1)
    masters = []
    for body in _sections(text, "capacitance"):
        masters += body.split()
    return {"window": np.array(window), "layer_z": np.array(layer_z), "layer_eps": np.array(layer_eps),
            "media_boxes": np.array(media_box).reshape(-1, 6), "media_eps": np.array(media_eps),
            "boxes": np.array(boxes).reshape(-1, 6), "ids": np.array(ids, np.int32), "names": names,
            "masters": masters}


def read(path: str) -> dict:
    with open(path) as f:
        return parse(f.read())


def read_reference_total(dspf_text: str):
    """(net, total capacitance in F) from the `*|NET name value` line of a
    reference .dspf result (testcases/cap3d/ref/case*.dspf)."""
    m = re.search(r"^\*\|NET\s+(\S+)\s+(\S+)", dspf_text, flags=re.M)
    if m:
        return m.group(1), float(m.group(2))
    raise ValueError("no *|NET line")


def master_capacitance(server, structure: dict, master: str, walks=20000, threads=16, concurrent=256, N=23, seed=1,
                       stop_frac=1e-3):
    """Total capacitance of `master` (F) and its standard error, with the
    walker on `server`; also the coupling row by conductor name."""
    if master not in structure["names"]:
        raise KeyError(f"master net {master!r} not among the conductors")
    k = structure["names"].index(master)
    scale = float(np.min(structure["boxes"][:, 3:] - structure["boxes"][:, :3]))
    cap, err, st = server.capacitance(structure, k, walks=walks, threads=threads, concurrent=concurrent, N=N,
                                      seed=seed, stop_dist=stop_frac * scale)
    row = {n: float(c) * EPS0_F_PER_UM for n, c in zip(structure["names"], cap)}
    return float(cap[k]) * EPS0_F_PER_UM, float(err[k]) * EPS0_F_PER_UM, row, st
