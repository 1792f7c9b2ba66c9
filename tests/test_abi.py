"""CPU checks of the C-ABI library: it loads and exports every symbol the
header declares, and argument errors come back as codes (no GPU needed)."""
import ctypes

import pytest


def test_library_exports_header_symbols():
    from deeprwcap_amd import _lib
    L = _lib.lib()
    declared = _lib.header_symbols()
    assert len(declared) >= 10
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) >= set(declared)


def test_argument_errors_without_gpu():
    from deeprwcap_amd import _lib
    L = _lib.lib()
    assert L.drw_abi_version() == 1
    rc = L.drw_fdm_solve(-1, None, None, None, None, None, None, None, 1e-6, None, None, None)
    assert rc == -2 and b"B < 0" in L.drw_last_error()
    rc = L.drw_fdm_solve(4, None, None, None, None, None, None, None, 1e-6, None, None, None)
    assert rc == -2 and b"null" in L.drw_last_error()
    rc = L.drw_solver_create(7, 0, ctypes.byref(ctypes.c_void_p()))
    assert rc == -2
    assert L.drw_solver_output_width(None) == -1
