"""Python mirror of the reference dnnsolver classes over the C ABI.

Mirrors inference_cpp/src/dnn_solver.h:43-149 (DNNSolver, DNNSolverPoisson,
DNNSolverGrad): same names, argument meaning and error behaviour.
``sample`` takes flat float lists/arrays of B*30 layer values and B*105
cuboid values and returns B*3 (Poisson) or B*5 (gradient) floats.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np

from . import _lib
from ._lib import DrwError, check

N = 23
NN = N * N
NNN = N * N * N
MAX_SAMPLES = 2048
MAX_CUBOIDS = 15
MAX_LAYERS = 15
PADDING_VALUE = -9999

POISSON, GRADIENT = 0, 1
BACKEND_NN, BACKEND_FDM = 0, 1


class DNNSolver:
    """Base: owns a C-ABI solver handle on one device (dnn_solver.cpp:13-38)."""

    kind = POISSON

    def __init__(self, device_idx: int = 0):
        h = ctypes.c_void_p()
        check(_lib.lib().drw_solver_create(self.kind, device_idx, ctypes.byref(h)), "drw_solver_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().drw_solver_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def loadModels(self, verbose: bool = False, models_txt: str = "models.txt"):
        check(_lib.lib().drw_solver_load_models(self._h, models_txt.encode()), "loadModels")
        if verbose:
            print(f"\tModels loaded from {models_txt}")

    def use_fdm(self, n: int = 23, tol: float = 1e-6):
        """Route queries through the batched FDM solve (the fallback and accuracy reference)."""
        check(_lib.lib().drw_solver_use_fdm(self._h, int(n), float(tol)), "use_fdm")

    def set_seed(self, seed: int, counter: int = 0):
        check(_lib.lib().drw_solver_set_seed(self._h, int(seed), int(counter)), "set_seed")

    @property
    def counter(self) -> int:
        c = ctypes.c_uint64()
        check(_lib.lib().drw_solver_get_counter(self._h, ctypes.byref(c)), "get_counter")
        return c.value

    @property
    def width(self) -> int:
        return _lib.lib().drw_solver_output_width(self._h)

    def _sample(self, layerStructures, cuboidStructures, axis):
        L = np.ascontiguousarray(layerStructures, dtype=np.float32).reshape(-1)
        C = np.ascontiguousarray(cuboidStructures, dtype=np.float32).reshape(-1)
        B = L.size // (MAX_LAYERS * 2)
        if C.size != B * MAX_CUBOIDS * 7:
            raise DrwError(f"cuboid array holds {C.size} floats, expected {B * MAX_CUBOIDS * 7}")
        A = None if axis is None else np.ascontiguousarray(axis, dtype=np.int32).reshape(-1)
        if A is not None and A.size != B:
            raise DrwError(f"axis array holds {A.size} entries, expected {B}")
        out = np.empty(B * self.width, dtype=np.float32)
        check(_lib.lib().drw_solver_sample(self._h, B, L.ctypes.data, C.ctypes.data,
                                           None if A is None else A.ctypes.data, out.ctypes.data), "sample")
        return out

    def sample_device(self, B, layers, cuboids, axis, out, stream=None):
        """Device-pointer variant (torch CUDA tensors), enqueued on `stream`."""
        check(_lib.lib().drw_solver_sample_device(self._h, int(B), _lib.ptr(layers), _lib.ptr(cuboids),
                                                  _lib.ptr(axis), _lib.ptr(out),
                                                  None if stream is None else _lib.stream_handle(stream)),
              "sample_device")


class DNNSolverPoisson(DNNSolver):
    """Poisson transition query (dnn_solver_poisson.cpp)."""

    kind = POISSON

    def sample(self, layerStructures, cuboidStructures, axis=None):
        # the axis is ignored for Green's-function sampling (dnn_solver_poisson.cpp:65-84)
        if axis is not None:
            warnings.warn("Axis parameter provided to Green's function solver but will be ignored.")
        return self._sample(layerStructures, cuboidStructures, None)


class DNNSolverGrad(DNNSolver):
    """Gradient transition query (dnn_solver_gradient.cpp)."""

    kind = GRADIENT

    def sample(self, layerStructures, cuboidStructures, axis=None):
        if axis is None:
            raise RuntimeError("Gradient solver requires axis parameter")  # dnn_solver_gradient.cpp:40-47
        return self._sample(layerStructures, cuboidStructures, axis)
