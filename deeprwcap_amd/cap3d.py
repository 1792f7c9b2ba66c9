"""Reader for the cap3d structure format of the reference's test cases
(testcases/cap3d/*.cap3d; semantics of the fields as the reference's
executable/run_script.py:32-76 drives them) and the end-to-end capacitance
harness around the walker.

Recognised sections:

* ``<window>`` v1(...) v2(...): the simulation box (potential 0 outside);
* ``<plate_medium>`` diel, z_top: planar dielectric layers, each from the
  previous z_top (or -inf) up to its own z_top;
* ``<medium>`` with ``<block>``s and diel: dielectric boxes (they override the
  planar layers where they lie, and a later box overrides an earlier one:
  in the test cases inner boxes are always listed after the boxes around
  them, never before);
* ``<conductor>`` name with ``<block>``s: conductor boxes;
* ``<task>`` ``<capacitance>`` names: the master nets;
* ``<layer>``, ``<parameter>``: read and ignored.

Conductor names of the form ``<k>__<net>`` are sections of net ``<net>``
(the test cases split ``signalGND`` into hundreds of them), and one net may
also be listed as several ``<conductor>`` sections of the same name; the
walker treats every net as one conductor.

Field-solver results (``ref/*.dspf``, ``ref/case10.spice``) are read into a
capacitance matrix {master: {net: F}} and compared with the walker's matrix
by the reference's metric (executable/run_script.py:32-76, 179-197).

A block is basepoint(x,y,z) + v1 + v2 + hvector; the test cases only use
axis-aligned edge vectors, which is required here (other blocks raise).
Corners are rounded to 1e-9 um so that faces given by different sums
coincide exactly.
Coordinates are in micrometres; capacitances come out in farads.
"""
from __future__ import annotations

import gzip
import re

import numpy as np

EPS0_F_PER_UM = 8.8541878128e-18  # vacuum permittivity, F/um
SNAP_DIGITS = 9  # block corners are rounded to 1e-9 um

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
    # a corner is a sum (0.076 + 0.041 != 0.117 in binary): snap to 1e-9 um,
    # so that faces meant to coincide do, without slivers of 1e-17 um
    return [*np.round(lo, SNAP_DIGITS), *np.round(hi, SNAP_DIGITS)]


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


def read_text(path: str) -> str:
    """Text of a file, gunzipped when its name ends in .gz."""
    if path.endswith(".gz"):
        with gzip.open(path, "rt") as f:
            return f.read()
    with open(path) as f:
        return f.read()


def read(path: str) -> dict:
    return parse(read_text(path))


def net_of(conductor_name: str) -> str:
    """Net of a conductor section: ``<k>__<net>`` -> ``<net>``, else the name."""
    head, sep, tail = conductor_name.partition("__")
    return tail if sep and head.isdigit() and tail else conductor_name


def nets(structure: dict):
    """(net names in order of first appearance, net index of every box)."""
    order, index = [], {}
    for name in structure["names"]:
        n = net_of(name)
        if n not in index:
            index[n] = len(order)
            order.append(n)
    sec = np.array([index[net_of(n)] for n in structure["names"]], np.int32)
    return order, sec[structure["ids"]] if len(structure["ids"]) else np.zeros(0, np.int32)


def _set_coupling(mat: dict, a: str, b: str, value: float):
    # a coupling listed in both rows is the mean of the two entries, stored in both
    if a in mat and b in mat[a]:
        value = (mat[a][b] + value) / 2
    mat[a][b] = value
    if b in mat:
        mat[b][a] = value


def read_reference_matrix(text: str) -> dict:
    """Capacitance matrix {master: {net: F}} of a field-solver .dspf result:
    ``*|NET <master> <total>`` opens a row whose diagonal is the total;
    ``C<k> <master> <net> <value>`` lines are couplings, net ``0`` meaning
    GROUND (parse_dspf, executable/run_script.py:56-76)."""
    mat = {}
    lines = text.splitlines()
    for line in lines:
        if line.startswith("*|NET "):
            t = line.split()
            mat[t[1]] = {t[1]: float(t[2])}
    for line in lines:
        if line.startswith("C"):
            t = line.split()
            _set_coupling(mat, t[1], "GROUND" if t[2] == "0" else t[2], float(t[3]))
    return mat


def _spice_value(tok: str) -> float:
    if tok.endswith("pF"):
        return float(tok[:-2]) * 1e-12
    if tok.endswith("fF"):
        return float(tok[:-2]) * 1e-15
    return float(tok)


def read_reference_spice(text: str) -> dict:
    """Capacitance matrix of a QuickCap .spice result: ``** <net> <value>
    ... (total)`` rows and ``Cp<k> <net> <net> <value>`` couplings, values in
    F, fF or pF (parse_spice, executable/run_script.py:24-54)."""
    mat = {}
    lines = text.splitlines()
    for line in lines:
        if line.startswith("** ") and "(total)" in line:
            t = line.split()
            mat[t[1]] = {t[1]: _spice_value(t[2])}
    for line in lines:
        if line.startswith("Cp"):
            t = line.split()
            _set_coupling(mat, t[1], "GROUND" if t[2] == "0" else t[2], _spice_value(t[3]))
    return mat


def read_reference(path: str) -> dict:
    """Reference matrix from a .dspf or .spice file (optionally .gz)."""
    text = read_text(path)
    return read_reference_spice(text) if ".spice" in path else read_reference_matrix(text)


def matrix_error(mat: dict, ref: dict):
    """(matrix error, {master: self-capacitance error}) of `mat` against
    `ref` (compute_error, executable/run_script.py:179-197): over the rows of
    `ref` whose master `mat` also holds, sum |C - C_ref| over the reference
    row's entries (diagonal included; a net missing from `mat` counts as 0)
    divided by the sum of |C_ref|; the self error of a row is
    |C_mm / C_ref,mm - 1|."""
    if mat is None or ref is None:
        return float("nan"), {}
    abserr = abssum = 0.0
    self_err = {}
    for m, row in ref.items():
        if m not in mat or m not in mat[m]:
            continue
        self_err[m] = abs(mat[m][m] / row[m] - 1)
        for n, c in row.items():
            abserr += abs(mat[m].get(n, 0.0) - c)
            abssum += abs(c)
    return (abserr / abssum if abssum else 0.0), self_err


def read_reference_total(dspf_text: str):
    """(net, total capacitance in F) from the `*|NET name value` line of a
    reference .dspf result (testcases/cap3d/ref/case*.dspf)."""
    m = re.search(r"^\*\|NET\s+(\S+)\s+(\S+)", dspf_text, flags=re.M)
    if m:
        return m.group(1), float(m.group(2))
    raise ValueError("no *|NET line")


def master_row(server, structure: dict, master: str, walks=20000, threads=16, concurrent=256, N=23, seed=1,
               stop_frac=1e-3):
    """Capacitance row of net `master` with the walker on `server`, as the
    reference's walker output reads it (executable/run_script.py:78-105:
    magnitudes by net name, the diagonal being the total): ({net: F},
    {net: standard error in F}, walker statistics)."""
    order, ids = nets(structure)
    if master not in order:
        raise KeyError(f"master net {master!r} not among the conductors")
    k = order.index(master)
    geo = dict(structure, ids=ids)
    scale = float(np.min(structure["boxes"][:, 3:] - structure["boxes"][:, :3]))
    cap, err, st = server.capacitance(geo, k, walks=walks, threads=threads, concurrent=concurrent, N=N, seed=seed,
                                      stop_dist=stop_frac * scale)
    row = {n: abs(float(c)) * EPS0_F_PER_UM for n, c in zip(order, cap)}
    row_err = {n: float(e) * EPS0_F_PER_UM for n, e in zip(order, err)}
    return row, row_err, st


def master_capacitance(server, structure: dict, master: str, **kw):
    """Total capacitance of `master` (F), its standard error, the row by net
    name and the walker statistics."""
    row, row_err, st = master_row(server, structure, master, **kw)
    return row[master], row_err[master], row, st
