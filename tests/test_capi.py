"""The C-ABI library loads and exports every entry point its header declares."""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "pdds_engine.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pdds_[a-z_0-9]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    from pdds_amd import _capi
    L = _capi.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding covers all of them
    assert set(names) <= set(_capi.EXPORTS)


def test_status_strings_and_defaults():
    from pdds_amd import _capi as A
    L = A.lib()
    for s in range(8):
        assert L.pdds_status_string(s)
    r = A.Run()
    L.pdds_run_defaults(C.byref(r))
    assert r.beta == -0.5826 and r.quirk_last_side == 1 and r.precision == A.FP64 and r.plain == 0


def test_invalid_arguments_are_rejected_without_a_device():
    from pdds_amd import _capi as A
    L = A.lib()
    h = C.c_void_p()
    fac = np.array([1.5, 1.27, 1.2, 1.13])
    assert L.pdds_mesh_create(0.0, 0.0, -1.0, 1.0, 4, 8, A.ptr(fac), C.byref(h)) == A.ERR_INVALID
    assert L.pdds_mesh_create(0.0, 0.0, 1.0, 1.0, 1, 8, A.ptr(fac), C.byref(h)) == A.ERR_INVALID
    bad = np.array([1.5, -1.0, 1.2, 1.13])
    assert L.pdds_mesh_create(0.0, 0.0, 1.0, 1.0, 4, 8, A.ptr(bad), C.byref(h)) == A.ERR_INVALID
    assert L.pdds_paths(None, None, None, 0.0, 0.0, None, 0, 0, 0, 0, *([None] * 12), None) == A.ERR_INVALID
    assert L.pdds_solve(None, 0, None, None, None, None, None, 0.0, 0, None, None) == A.ERR_INVALID
