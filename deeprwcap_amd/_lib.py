"""ctypes binding of libdnnsolver.so (the C ABI in include/dnnsolver_c.h).

The library must have been built (``python -m deeprwcap_amd.build``); there is
no CPU fallback: every entry point runs the HIP kernels or raises.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdnnsolver.so")
HEADER = os.path.join(HERE, "..", "include", "dnnsolver_c.h")

_c_int, _c_dbl, _c_u64, _c_vp = ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p

# name -> (restype, argtypes); pointers are passed as integers (c_void_p)
SIGNATURES = {
    "drw_last_error": (ctypes.c_char_p, []),
    "drw_abi_version": (_c_int, []),
    "drw_fdm_solve": (_c_int, [_c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp, _c_vp,
                               _c_vp]),
    "drw_voxelize": (_c_int, [_c_int, _c_vp, _c_vp, _c_vp, _c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp]),
    "drw_sample_flat": (_c_int, [_c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_u64, _c_u64, _c_vp, _c_vp]),
    "drw_solver_create": (_c_int, [_c_int, _c_int, ctypes.POINTER(_c_vp)]),
    "drw_solver_destroy": (None, [_c_vp]),
    "drw_solver_load_models": (_c_int, [_c_vp, ctypes.c_char_p]),
    "drw_solver_use_fdm": (_c_int, [_c_vp, _c_int, _c_dbl]),
    "drw_solver_backend": (_c_int, [_c_vp]),
    "drw_solver_set_seed": (_c_int, [_c_vp, _c_u64, _c_u64]),
    "drw_solver_get_counter": (_c_int, [_c_vp, ctypes.POINTER(_c_u64)]),
    "drw_solver_output_width": (_c_int, [_c_vp]),
    "drw_solver_set_fdm_limits": (_c_int, [_c_vp, _c_int, _c_int]),
    "drw_solver_failures": (_c_int, [_c_vp, ctypes.POINTER(_c_u64), ctypes.POINTER(_c_u64), ctypes.POINTER(_c_u64)]),
    "drw_set_launch_shape": (_c_int, [_c_int, _c_int, _c_int]),
    "drw_get_launch_shape": (_c_int, [_c_int, _c_int]),
    "drw_solver_set_timing": (_c_int, [_c_vp, _c_int]),
    "drw_solver_stage_times": (_c_int, [_c_vp, _c_vp, _c_vp, _c_int]),
    "drw_solver_sample": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp]),
    "drw_solver_sample_device": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "drw_solver_sample_device_noalloc": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "drw_solver_sample_ids": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "drw_solver_graph_create": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, ctypes.POINTER(_c_vp)]),
    "drw_graph_launch": (_c_int, [_c_vp, _c_vp]),
    "drw_graph_counter": (_c_int, [_c_vp, ctypes.POINTER(_c_u64)]),
    "drw_graph_destroy": (None, [_c_vp]),
    "drw_solver_sample_ex": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                                      _c_vp]),
    "drw_server_create": (_c_int, [_c_int, ctypes.c_char_p, _c_vp, ctypes.POINTER(_c_vp)]),
    "drw_server_submit": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_vp, _c_vp, _c_u64, ctypes.POINTER(_c_u64)]),
    "drw_server_wait": (_c_int, [_c_vp, _c_u64, _c_vp, _c_int]),
    "drw_server_submit_many": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "drw_server_wait_many": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_int]),
    "drw_server_stats_get": (_c_int, [_c_vp, _c_vp, _c_int]),
    "drw_server_resident_stats": (_c_int, [_c_vp, _c_vp]),
    "drw_server_destroy": (None, [_c_vp]),
    "drw_walk_capacitance": (_c_int, [_c_vp, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp]),
    "drw_server_run_load": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_int, _c_vp]),
    "drw_solver_eval_kernels": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
}

_LIB = None


class DrwError(RuntimeError):
    pass


def header_symbols():
    """Function names declared in include/dnnsolver_c.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(drw_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise DrwError(f"{LIB_PATH} is missing: build it with `python -m deeprwcap_amd.build`")
        # torch ships its own HIP runtime (same soname, libamdhip64.so.7) and
        # asks for it by file name, so it only shares a runtime with this
        # library when it is loaded first; two runtimes in one process leave
        # torch without a GPU.  Load torch before the library.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().drw_last_error().decode(errors="replace")
        raise DrwError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> int | None:
    """Device (or host) address of a torch tensor / numpy array, None for None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_handle(stream=None) -> int | None:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
