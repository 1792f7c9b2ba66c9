"""Network weight files (.drwnet) and the models.txt model list.

.drwnet (little endian): b"DRWNET01", int32 count, then per tensor
int32 name length, name, int32 ndim, int32 dims[ndim], float32 data.
Tensor names are those of the reference state dicts
(training_pytorch/src/architecture.py); the library folds batch norm and
packs MFMA fragments when it loads them.

models.txt follows the reference list (inference_cpp/models.txt:4-8):
index 0 PoissonSelector, 1 PoissonPredictor, 2 GradientSelectorWeight,
3 Gradient2Predictor, 4 Gradient1Predictor.
"""
from __future__ import annotations

import os
import struct

import numpy as np

MODEL_ORDER = ["PoissonSelector", "PoissonPredictor", "GradientSelectorWeight", "Gradient2Predictor",
               "Gradient1Predictor"]
MAGIC = b"DRWNET01"

FACE_SOLVER = (16, [(16, 16, 1), (16, 8, 1), (8, 4, 2), (4, 2, 3)])
FACE_SOLVER_PLUS = (32, [(32, 32, 1), (32, 24, 1), (24, 24, 1), (24, 16, 1), (16, 16, 2), (16, 8, 2), (8, 4, 3)])


def write_drwnet(path: str, state: dict) -> None:
    with open(path, "wb") as f:
        f.write(MAGIC)
        items = [(k, np.ascontiguousarray(v, dtype=np.float32)) for k, v in state.items()]
        f.write(struct.pack("<i", len(items)))
        for k, v in items:
            kb = k.encode()
            f.write(struct.pack("<i", len(kb)))
            f.write(kb)
            f.write(struct.pack("<i", v.ndim))
            f.write(struct.pack(f"<{v.ndim}i", *v.shape))
            f.write(v.tobytes())


def read_drwnet(path: str) -> dict:
    out = {}
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path} is not a DRWNET01 file")
        (n,) = struct.unpack("<i", f.read(4))
        for _ in range(n):
            (ln,) = struct.unpack("<i", f.read(4))
            name = f.read(ln).decode()
            (nd,) = struct.unpack("<i", f.read(4))
            shape = struct.unpack(f"<{nd}i", f.read(4 * nd)) if nd else ()
            cnt = int(np.prod(shape)) if nd else 1
            out[name] = np.frombuffer(f.read(4 * cnt), dtype=np.float32).reshape(shape).copy()
    return out


def write_models_txt(directory: str, files: dict) -> str:
    """files: model name -> .drwnet path (relative to directory or absolute)."""
    path = os.path.join(directory, "models.txt")
    with open(path, "w") as f:
        f.write("# Model configuration file\n# Format: path # description\n\n")
        for name in MODEL_ORDER:
            f.write(f"{files[name]} # {name}\n")
    return path


def _bn(prefix, C, rng):
    return {prefix + ".weight": 0.5 + rng.random(C), prefix + ".bias": 0.1 * rng.standard_normal(C),
            prefix + ".running_mean": 0.2 * rng.standard_normal(C), prefix + ".running_var": 0.5 + rng.random(C)}


def _face_shapes(c0, blocks, rng):
    s = {"solver.z_weight.0.weight": rng.standard_normal((c0, 25, 1, 1)) / 5.0}
    s.update(_bn("solver.z_weight.1", c0, rng))
    for i, (ci, co, _) in enumerate(blocks):
        k = f"solver.blocks.{i}.block."
        s[k + "0.weight"] = rng.standard_normal((ci, 1, 3, 3)) / 3.0
        s.update(_bn(k + "1", ci, rng))
        s[k + "3.weight"] = rng.standard_normal((co, ci, 1, 1)) / np.sqrt(ci)
        s.update(_bn(k + "4", co, rng))
    clast = blocks[-1][1]
    s["solver.head.weight"] = rng.standard_normal((1, clast, 1, 1)) / np.sqrt(clast)
    s["solver.head.bias"] = 0.1 * rng.standard_normal(1)
    return s


def _sel_shapes(weighted, rng):
    ch = [1, 8, 16, 64] if weighted else [1, 4, 8, 16]
    s = {}
    for l in range(3):
        ci, co = ch[l], ch[l + 1]
        s[f"backbone.{3 * l}.weight"] = rng.standard_normal((co, ci, 3, 3, 3)) / np.sqrt(27 * ci)
        s[f"backbone.{3 * l}.bias"] = 0.1 * rng.standard_normal(co)
        s.update(_bn(f"backbone.{3 * l + 1}", co, rng))
    dims = [512, 256, 128, 7] if weighted else [128, 64, 6]
    for j in range(len(dims) - 1):
        s[f"head.{2 * j + 1}.weight"] = rng.standard_normal((dims[j + 1], dims[j])) / np.sqrt(dims[j])
        s[f"head.{2 * j + 1}.bias"] = 0.1 * rng.standard_normal(dims[j + 1])
    return s


def random_state(name: str, seed: int = 0) -> dict:
    """Seeded random weights (and batch-norm statistics) for one network."""
    rng = np.random.default_rng(seed)
    if name == "PoissonSelector":
        s = _sel_shapes(False, rng)
    elif name == "GradientSelectorWeight":
        s = _sel_shapes(True, rng)
    elif name == "Gradient2Predictor":
        s = _face_shapes(*FACE_SOLVER_PLUS, rng)
    else:
        s = _face_shapes(*FACE_SOLVER, rng)
    return {k: np.asarray(v, np.float32) for k, v in s.items()}


def state_from_npz(d, name: str) -> dict:
    pre = name + "/"
    return {k[len(pre):]: np.asarray(d[k], np.float32) for k in d.files
            if k.startswith(pre) and not k.endswith("/output")}


def export_model_dir(directory: str, states: dict) -> str:
    """Write one .drwnet per network plus models.txt; returns models.txt."""
    os.makedirs(directory, exist_ok=True)
    files = {}
    for name in MODEL_ORDER:
        fn = f"{name}.drwnet"
        write_drwnet(os.path.join(directory, fn), states[name])
        files[name] = fn
    return write_models_txt(directory, files)
