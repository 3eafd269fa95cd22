"""Kuramoto oscillators on the device (host mirror of simuse_kura_*).

Same coupling as the reference's kuramoto_step_sparse
(crates/SiMuse/crates/pillars/src/kernels/phase/sparse.rs:40-97):
c_i = sum_j W_ij sin(theta_j - theta_i), theta += dt (omega + c), wrapped to
[-pi, pi); ``method="rk4"`` integrates the same ODE with classic RK4.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class KuramotoGraph:
    def __init__(self, row_ptr, col, w, omega, theta0, dt: float, method: str = "rk4"):
        rp = np.ascontiguousarray(row_ptr, np.uint32)
        cc = np.ascontiguousarray(col, np.uint32)
        ww = np.ascontiguousarray(w, np.float32)
        om = np.ascontiguousarray(omega, np.float32)
        th = np.ascontiguousarray(theta0, np.float32)
        self.n = int(om.size)
        self.n_edges = int(cc.size)
        m = _lib.KURA_RK4 if method == "rk4" else _lib.KURA_EULER
        h = C.c_void_p()
        check(lib().simuse_kura_create(_p(rp, C.c_uint32), _p(cc, C.c_uint32), _p(ww, C.c_float),
                                       _p(om, C.c_float), _p(th, C.c_float), self.n, float(dt), m, C.byref(h)))
        self._h = h

    @classmethod
    def generated(cls, desc, method: str = "rk4") -> "KuramotoGraph":
        """A modular small-world graph built on the device from a KuraGenDesc
        (simuse_kura_create_generated)."""
        from .server import _S
        L = _S()
        self = cls.__new__(cls)
        h = C.c_void_p()
        m = _lib.KURA_RK4 if method == "rk4" else _lib.KURA_EULER
        check(L.simuse_kura_create_generated(C.byref(desc), m, C.byref(h)))
        self._h = h
        n, ne = C.c_uint32(), C.c_uint64()
        check(L.simuse_kura_info(h, C.byref(n), C.byref(ne)))
        self.n, self.n_edges = int(n.value), int(ne.value)
        return self

    def download_graph(self):
        """(row_ptr, col, w, omega) of the graph."""
        from .server import _S
        rp = np.zeros(self.n + 1, np.uint32)
        col = np.zeros(max(1, self.n_edges), np.uint32)
        w = np.zeros(max(1, self.n_edges), np.float32)
        om = np.zeros(self.n, np.float32)
        check(_S().simuse_kura_download_graph(self._h, _p(rp, C.c_uint32), _p(col, C.c_uint32), _p(w, C.c_float),
                                              _p(om, C.c_float)))
        return rp, col[: self.n_edges], w[: self.n_edges], om

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().simuse_kura_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_group(self, lanes: int):
        check(lib().simuse_kura_set_group(self._h, int(lanes)))

    def step_async(self, n: int):
        check(lib().simuse_kura_step_async(self._h, int(n)))

    def sync(self):
        check(lib().simuse_kura_sync(self._h))

    def run(self, n: int):
        self.step_async(n)
        self.sync()

    def phases(self) -> np.ndarray:
        out = np.zeros(self.n, np.float64)
        check(lib().simuse_kura_download_phases(self._h, _p(out, C.c_double), self.n))
        return out

    def order_parameter(self) -> float:
        r = C.c_double()
        check(lib().simuse_kura_order_parameter(self._h, C.byref(r)))
        return float(r.value)

    def phase_coherence(self, amplitudes) -> float:
        """|sum a e^{i theta}| / sum a (measure/mod.rs:76-106)."""
        a = np.ascontiguousarray(amplitudes, np.float32)
        if a.size != self.n:
            raise ValueError("one amplitude per oscillator")
        r = C.c_double()
        check(lib().simuse_kura_phase_coherence(self._h, _p(a, C.c_float), C.byref(r)))
        return float(r.value)

    def step_counter(self) -> int:
        s = C.c_uint64()
        check(lib().simuse_kura_step_counter(self._h, C.byref(s)))
        return int(s.value)

    def set_profiling(self, on: bool):
        check(lib().simuse_kura_set_profiling(self._h, 1 if on else 0))

    def step_ms(self) -> float:
        out = (C.c_float * 1)()
        check(lib().simuse_kura_kernel_times(self._h, out, 1))
        return float(out[0])
