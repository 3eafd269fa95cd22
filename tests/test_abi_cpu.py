"""The C-ABI library loads, exports every symbol include/*.h declares, and
rejects malformed networks on the host (CPU only: no kernel is launched)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import simuse
from simuse import _lib
from simuse.models import izh_pop, lif_pop
from simuse.network import PopDesc, _descs

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(REPO, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            names += re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?[\s\*])(simuse_\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _create(pops, rp, col, w, d, dt=0.0):
    descs = _descs(pops)
    h = C.c_void_p()
    rp = np.ascontiguousarray(rp, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    w = np.ascontiguousarray(w, np.float32)
    d = np.ascontiguousarray(d, np.uint8)
    rc = _lib.lib().simuse_net_create(descs, len(pops), rp.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      col.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      w.ctypes.data_as(C.POINTER(C.c_float)),
                                      d.ctypes.data_as(C.POINTER(C.c_uint8)), col.size, dt, C.byref(h))
    if rc == 0:
        _lib.lib().simuse_net_destroy(h)
    return rc, _lib.lib().simuse_last_error().decode()


def _pop(n=4):
    return lif_pop(n, 0.1, 0.05, -65.0, -50.0, -65.0, 2.0, -65.0)


@pytest.mark.parametrize("case,msg", [
    ("rp0", "start at 0"),
    ("mono", "non-decreasing"),
    ("end", "end at"),
    ("col", "col index"),
    ("w", "finite"),
    ("delay", "MAX_DELAY_SLOTS"),
    ("dt", "differs"),
])
def test_create_rejects_malformed_csr(case, msg):
    # SparseGraph::from_csr validation (pillars/src/api/net/graph.rs:217-280)
    rp = np.array([0, 1, 2, 3, 4], np.uint32)
    col = np.array([1, 2, 3, 0], np.uint32)
    w = np.ones(4, np.float32)
    d = np.zeros(4, np.uint8)
    dt = 0.0
    if case == "rp0":
        rp = np.array([1, 1, 2, 3, 4], np.uint32)
    elif case == "mono":
        rp = np.array([0, 2, 1, 3, 4], np.uint32)
    elif case == "end":
        rp = np.array([0, 1, 2, 3, 3], np.uint32)
    elif case == "col":
        col[2] = 9
    elif case == "w":
        w[1] = np.inf
    elif case == "delay":
        d[0] = 32
    elif case == "dt":
        dt = 0.5
    rc, err = _create([_pop()], rp, col, w, d, dt)
    assert rc == _lib.E_INVALID
    assert msg in err


def test_create_rejects_bad_populations():
    p = _pop()
    desc = (PopDesc * 1)()
    desc[0] = p.desc()
    desc[0].model = 7
    h = C.c_void_p()
    rp = np.zeros(5, np.uint32)
    rc = _lib.lib().simuse_net_create(desc, 1, rp.ctypes.data_as(C.POINTER(C.c_uint32)), None, None, None, 0,
                                      0.0, C.byref(h))
    assert rc == _lib.E_UNSUPPORTED
    desc[0].model = 0
    desc[0].ns = 3
    rc = _lib.lib().simuse_net_create(desc, 1, rp.ctypes.data_as(C.POINTER(C.c_uint32)), None, None, None, 0,
                                      0.0, C.byref(h))
    assert rc == _lib.E_INVALID and "needs 2 states" in _lib.lib().simuse_last_error().decode()
    rc, err = _create([lif_pop(0, 0.1, 0.05, -65, -50, -65, 2, -65)], np.zeros(1, np.uint32), [], [], [])
    assert rc == _lib.E_INVALID and "no neurons" in err


@pytest.mark.skipif(simuse.device_count() > 0, reason="host without a GPU only")
def test_create_without_device_reports_hip_error():
    rc, err = _create([_pop()], np.zeros(5, np.uint32), [], [], [])
    assert rc == _lib.E_HIP and "device" in err


def test_null_handle_is_an_error_not_a_crash():
    L = _lib.lib()
    assert L.simuse_net_step(None) == _lib.E_INVALID
    L.simuse_net_destroy(None)


def test_population_api_rejects_bad_arguments_on_the_host():
    # simuse_pop_create validates before touching the device
    from simuse.model import _L
    L = _L()
    h = C.c_void_p()
    assert L.simuse_pop_create(7, 2, 6, 10, 16, C.byref(h)) == _lib.E_INVALID  # unknown model
    assert L.simuse_pop_create(simuse.MODEL_LIF, 3, 6, 10, 16, C.byref(h)) == _lib.E_INVALID  # ns
    assert L.simuse_pop_create(simuse.MODEL_LIF, 2, 7, 10, 16, C.byref(h)) == _lib.E_INVALID  # np
    assert L.simuse_pop_create(simuse.MODEL_LIF, 2, 6, 10, 8, C.byref(h)) == _lib.E_INVALID  # stride < n
    assert L.simuse_pop_create(simuse.MODEL_LIF, 2, 6, 0, 8, C.byref(h)) == _lib.E_INVALID  # empty
    assert L.simuse_pop_create(simuse.MODEL_LIF, 2, 6, 10, 16, None) == _lib.E_INVALID
    assert not h.value
    assert L.simuse_pop_upload(None, None, None, None) == _lib.E_INVALID
    assert L.simuse_pop_step_async(None, 0.1, 1) == _lib.E_INVALID
    assert L.simuse_pop_compact_spikes(None, None) == _lib.E_INVALID
    L.simuse_pop_destroy(None)
