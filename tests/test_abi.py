"""The C ABI (include/ztm_gfm.h) without a GPU: the library loads, exports every
declared entry point, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

from ztm_amd import gfm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ztm_gfm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ztm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for n in ("ztm_gf", "ztm_gf_async", "ztm_gfm_host", "ztm_plan_create",
              "ztm_plan_forward", "ztm_plan_adjoint"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gfm.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly the declared set
    assert sorted(gfm.SIGNATURES) == declared_functions()


def test_version_and_strerror():
    L = gfm.lib()
    assert L.ztm_version() == 100
    assert L.ztm_strerror(0) == b"ok"
    assert L.ztm_strerror(1) == b"invalid argument"
    assert L.ztm_strerror(99) == b"unknown error"


def test_bad_arguments_fail_before_any_device_call():
    L = gfm.lib()
    # reference ABI: bad geometry and null pointers
    assert L.ztm_gf(10, 0, 61, 33, 1, 2, None, None, None) == 1
    assert L.ztm_gf(10, 61, 61, 33, 1, 2, None, None, None) == 1
    assert b"null" in L.ztm_last_error()
    # GFM host body: rho of the wrong length (the dk=33/35 mismatch, GFM_GPU.cu:24)
    rho = np.zeros(61 * 61 * 69, np.float32)
    s = np.zeros(61 * 61, np.float32)
    g = np.zeros(61 * 61, np.float32)
    rc = L.ztm_gfm_host(10, 61, 61, 33, 1, 2, g.ctypes.data, rho.ctypes.data, rho.size,
                        s.ctypes.data)
    assert rc == 1
    assert b"ii*jj*kk" in L.ztm_last_error()
    # plans
    assert L.ztm_plan_set_tuning(None, 4, 0, 0, 0) == 1
    assert L.ztm_plan_forward(None, 0, None, None, None) == 1
    assert L.ztm_plan_num_cells(None) == -1
    with pytest.raises(gfm.ZtmError):
        gfm.check(1)


def test_plan_rejects_bad_problems_without_a_device():
    L = gfm.lib()
    h = ctypes.c_void_p()
    bad = gfm.CProblem(gfm.CGrid(0, 4, 4, 0, 0, 0, 1, 1, 1), 0, 1.0, 0.0)
    assert L.ztm_plan_create(ctypes.byref(h), ctypes.byref(bad), 1, None) == 1
    zero_dx = gfm.CProblem(gfm.CGrid(4, 4, 4, 0, 0, 0, 0.0, 1, 1), 0, 1.0, 0.0)
    assert L.ztm_plan_create(ctypes.byref(h), ctypes.byref(zero_dx), 1, None) == 1
    bad_formula = gfm.CProblem(gfm.CGrid(4, 4, 4, 0, 0, 0, 1, 1, 1), 9, 1.0, 0.0)
    assert L.ztm_plan_create(ctypes.byref(h), ctypes.byref(bad_formula), 1, None) == 1
    ok = gfm.CProblem(gfm.CGrid(4, 4, 4, 0, 0, 0, 1, 1, 1), 0, 1.0, 0.0)
    assert L.ztm_plan_create(ctypes.byref(h), ctypes.byref(ok), 5, None) == 1   # null stations
    assert L.ztm_plan_create(ctypes.byref(h), ctypes.byref(ok), -1, None) == 1


def test_batch_rejects_bad_problems_without_a_device():
    L = gfm.lib()
    h = ctypes.c_void_p()
    ok = gfm.CProblem(gfm.CGrid(4, 4, 4, 0, 0, 0, 1, 1, 1), 0, 1.0, 0.0)
    prism = gfm.CProblem(gfm.CGrid(4, 4, 4, 0, 0, 0, 1, 1, 1), 1, 1.0, 0.0)
    st = (ctypes.c_double * 3)(0.0, 0.0, 1.0)
    sp = (ctypes.c_void_p * 2)(ctypes.cast(st, ctypes.c_void_p), ctypes.cast(st, ctypes.c_void_p))
    probs = (gfm.CProblem * 2)(ok, prism)
    ns = (ctypes.c_int64 * 2)(1, 1)
    assert L.ztm_batch_create(ctypes.byref(h), 2, probs, ns, sp) == 1          # prism
    probs = (gfm.CProblem * 2)(ok, ok)
    ns0 = (ctypes.c_int64 * 2)(1, 0)
    assert L.ztm_batch_create(ctypes.byref(h), 2, probs, ns0, sp) == 1         # no stations
    assert L.ztm_batch_create(ctypes.byref(h), 0, probs, ns, sp) == 1          # empty batch
    assert L.ztm_batch_create(None, 2, probs, ns, sp) == 1
    assert h.value is None
    assert L.ztm_batch_forward(None, None, None, None) == 1
    assert L.ztm_batch_solve(None, None, None, 1, None, None, 1, None) == 1
