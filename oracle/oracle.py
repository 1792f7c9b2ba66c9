"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the CPU-baseline leg of bench.py, always as the
checker, never as the thing measured or shipped.

Every function names the reference file:line it restates (see oracle.c).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GFT, WVTZ, WVTX, WVTY = 0, 1, 2, 3
PAD = -9999.0
MAX_LAYERS = 15
MAX_CUBOIDS = 15


def build() -> None:
    """Compile liboracle.so (and oracle/_ref where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.orc_fdm_apply_L.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_fdm_diag.argtypes = [ctypes.c_int, dp, dp]
        L.orc_fdm_rhs.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp]
        L.orc_panel_cells.argtypes = [ctypes.c_int, ip]
        L.orc_fdm_cg.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.orc_fdm_cg.restype = ctypes.c_int
        L.orc_fdm_solve.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_fdm_solve.restype = ctypes.c_int
        L.orc_ggft_dense_A.argtypes = [ctypes.c_int, dp, dp]
        L.orc_center_eps.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.orc_center_eps.restype = ctypes.c_double
        L.orc_voxelize.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp, ctypes.c_int, fp, fp]
        L.orc_rotate_poisson.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp]
        L.orc_rotate_gradient.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp]
        L.orc_axis_rotate.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp]
        L.orc_post_process.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, fp]
        L.orc_locate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_query_uniforms.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, u32p]
        L.orc_cdf_draw_f32.argtypes = [fp, ctypes.c_int, ctypes.c_uint32]
        L.orc_cdf_draw_f32.restype = ctypes.c_int
        L.orc_cdf_draw_f64.argtypes = [dp, ctypes.c_int, ctypes.c_uint32]
        L.orc_cdf_draw_f64.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- FDM (GGFT)
def apply_L(eps, y):
    """(L y) with harmonic-mean faces, ggft_solver.cpp:447-518 (as L = A D)."""
    N = eps.shape[-1]
    eps, y = _f64(eps).reshape(-1), _f64(y).reshape(-1)
    out = np.empty_like(eps)
    lib().orc_fdm_apply_L(N, _p(eps, ctypes.c_double), _p(y, ctypes.c_double), _p(out, ctypes.c_double))
    return out


def rhs(eps, code):
    """Right-hand side e of ggft_solver.cpp:137-274 for odd and even N."""
    N = eps.shape[-1]
    eps = _f64(eps).reshape(-1)
    e = np.zeros_like(eps)
    lib().orc_fdm_rhs(N, _p(eps, ctypes.c_double), code, _p(e, ctypes.c_double))
    return e


def panel_cells(N):
    """Boundary-panel -> cell map in GGFT numbering (ggft_solver.h:65-77)."""
    out = np.empty(6 * N * N, dtype=np.int32)
    lib().orc_panel_cells(N, _p(out, ctypes.c_int))
    return out


def fdm_solve(eps, code=GFT, tol=1e-14, maxit=20000):
    """G (6N^2, float64) for one cube, ggft_solver.cpp:521-636 with G = 2 eps y.

    Returns (G, iterations, relative residual)."""
    N = eps.shape[-1]
    eps = _f64(eps).reshape(-1)
    G = np.empty(6 * N * N)
    rr = ctypes.c_double(0.0)
    it = lib().orc_fdm_solve(N, _p(eps, ctypes.c_double), code, tol, maxit, _p(G, ctypes.c_double),
                             ctypes.byref(rr))
    return G, it, rr.value


def ggft_dense_A(eps):
    """GGFT's own assembled matrix A (dense, tiny N only), ggft_solver.cpp:447-518."""
    N = eps.shape[-1]
    n = N ** 3
    eps = _f64(eps).reshape(-1)
    A = np.empty((n, n))
    lib().orc_ggft_dense_A(N, _p(eps, ctypes.c_double), _p(A, ctypes.c_double))
    return A


def center_eps(eps, code):
    """Centre permittivity of ggft_solver.cpp:649-671."""
    N = eps.shape[-1]
    eps = _f64(eps).reshape(-1)
    return lib().orc_center_eps(N, _p(eps, ctypes.c_double), code)


# ------------------------------------------------------------- voxelization
def voxelize(layers, cuboids, N, normalise=True):
    """fill_structure_kernel (cuda_kernels.cu:21-84) after the per-query
    normalisation of dnn_solver.cpp:274-285.  Returns (cube[B,N,N,N], max[B])."""
    layers, cuboids = _f32(layers), _f32(cuboids)
    B = layers.shape[0]
    out = np.empty((B, N, N, N), dtype=np.float32)
    mx = np.empty(B, dtype=np.float32)
    lib().orc_voxelize(B, N, _p(layers, ctypes.c_float), _p(cuboids, ctypes.c_float), int(normalise),
                       _p(out, ctypes.c_float), _p(mx, ctypes.c_float))
    return out, mx


# ------------------------------------------------------------------ remaps
def _cube_op(fn, N, arg, cube):
    cube = _f32(cube)
    out = np.empty_like(cube)
    fn(N, arg, _p(cube, ctypes.c_float), _p(out, ctypes.c_float))
    return out


def rotate_poisson(cube, face):
    """rotate_faces_poisson_kernel, cuda_kernels.cu:299-380."""
    return _cube_op(lib().orc_rotate_poisson, cube.shape[-1], face, cube)


def rotate_gradient(cube, face):
    """rotate_faces_gradient_kernel, cuda_kernels.cu:411-486."""
    return _cube_op(lib().orc_rotate_gradient, cube.shape[-1], face, cube)


def axis_rotate(cube, axis):
    """Axis pre-rotation of dnn_solver_gradient.cpp:228-238."""
    return _cube_op(lib().orc_axis_rotate, cube.shape[-1], axis, cube)


def post_process(face2d, face, axis):
    """post_process_gradient_kernel, cuda_kernels.cu:516-578."""
    N = face2d.shape[-1]
    face2d = _f32(face2d)
    out = np.empty_like(face2d)
    lib().orc_post_process(N, face, axis, _p(face2d, ctypes.c_float), _p(out, ctypes.c_float))
    return out


def locate(N, tile, face, axis=-1):
    """locate_index(_simple)_kernel, cuda_kernels.cu:122-264 (axis<0: simple)."""
    out = np.empty(3, dtype=np.float32)
    lib().orc_locate(N, int(tile), int(face), int(axis), _p(out, ctypes.c_float))
    return out


# ---------------------------------------------------------------- sampling
def philox(ctr, key):
    """Philox4x32-10 (Salmon et al., SC'11)."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.empty(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, ctypes.c_uint32), _p(k, ctypes.c_uint32), _p(o, ctypes.c_uint32))
    return o


def query_uniforms(seed, q, stream=0):
    o = np.empty(4, dtype=np.uint32)
    lib().orc_query_uniforms(ctypes.c_uint64(seed), ctypes.c_uint64(q), stream, _p(o, ctypes.c_uint32))
    return o


def cdf_draw(w, u):
    """Fixed-point inverse-CDF draw (integer cumulative sums)."""
    w = np.asarray(w)
    if w.dtype == np.float64:
        w = _f64(w)
        return lib().orc_cdf_draw_f64(_p(w, ctypes.c_double), w.size, int(u))
    w = _f32(w)
    return lib().orc_cdf_draw_f32(_p(w, ctypes.c_float), w.size, int(u))
