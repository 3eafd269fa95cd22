"""One population without synapses (simuse_pop_*): the mirror of
pillars_cuda::CompiledModel (crates/SiMuse/crates/pillars-cuda/src/
engine.rs:98-423).

States, params and the input current live on the device; each step advances
every neuron with the device-held input and the call's dt.  Host arrays are
column-major SoA with column stride `stride` (states[k * stride + i]).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import P, check, f32p, i32p, lib, u32p
from .network import N_PARAMS, N_STATE, _ptr, stride_for

_SIGS = {
    "simuse_pop_create": ([C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P)], C.c_int),
    "simuse_pop_destroy": ([P], None),
    "simuse_pop_upload": ([P, f32p, f32p, f32p], C.c_int),
    "simuse_pop_upload_i_in": ([P, f32p], C.c_int),
    "simuse_pop_step_async": ([P, C.c_float, C.c_uint32], C.c_int),
    "simuse_pop_sync": ([P], C.c_int),
    "simuse_pop_time_steps": ([P, C.c_float, C.c_uint32, C.POINTER(C.c_float)], C.c_int),
    "simuse_pop_step_device": ([P, C.c_float], C.c_int),
    "simuse_pop_compact_spikes": ([P, u32p], C.c_int),
    "simuse_pop_download_compact": ([P, i32p, C.c_uint32], C.c_int),
    "simuse_pop_download_spikes": ([P, i32p, C.c_uint64], C.c_int),
    "simuse_pop_download_states": ([P, f32p, C.c_uint64], C.c_int),
    "simuse_pop_step": ([P, f32p, f32p, f32p, i32p, C.c_float], C.c_int),
    "simuse_pop_run": ([P, f32p, f32p, f32p, i32p, C.c_float, C.c_uint32], C.c_int),
}

_bound = False


def _L():
    global _bound
    L = lib()
    if not _bound:
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _bound = True
    return L


def _f32(a, size, what):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    if a.size != size:
        raise ValueError(f"{what}: {a.size} values, expected {size}")
    return a


class CompiledModel:
    """CompiledModel::new(model, ns, np, n, stride); the model is a
    simuse.MODEL_* enum instead of an NVRTC source string."""

    def __init__(self, model: int, n: int, stride: int | None = None):
        self.model, self.n = int(model), int(n)
        self.ns, self.np = N_STATE[self.model], N_PARAMS[self.model]
        self.stride = stride_for(self.n) if stride is None else int(stride)
        h = P()
        check(_L().simuse_pop_create(self.model, self.ns, self.np, self.n, self.stride, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _L().simuse_pop_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, states, params, i_in) -> None:
        st = _f32(states, self.ns * self.stride, "states")
        pr = _f32(params, self.np * self.stride, "params")
        cur = _f32(i_in, self.n, "i_in")
        check(_L().simuse_pop_upload(self._h, _ptr(st, C.c_float), _ptr(pr, C.c_float), _ptr(cur, C.c_float)))

    def upload_i_in(self, i_in) -> None:
        cur = _f32(i_in, self.n, "i_in")
        check(_L().simuse_pop_upload_i_in(self._h, _ptr(cur, C.c_float)))

    def step_graph(self, dt: float, n_steps: int = 1) -> None:
        check(_L().simuse_pop_step_async(self._h, float(dt), int(n_steps)))

    def sync(self) -> None:
        check(_L().simuse_pop_sync(self._h))

    def time_steps(self, dt: float, n_steps: int) -> float:
        """Device milliseconds of n_steps steps (queued behind a spacer)."""
        ms = C.c_float()
        check(_L().simuse_pop_time_steps(self._h, float(dt), int(n_steps), C.byref(ms)))
        return float(ms.value)

    def step_device(self, dt: float) -> None:
        check(_L().simuse_pop_step_device(self._h, float(dt)))

    def compact_spikes(self) -> int:
        c = C.c_uint32()
        check(_L().simuse_pop_compact_spikes(self._h, C.byref(c)))
        return int(c.value)

    def download_compact(self, count: int) -> np.ndarray:
        out = np.zeros(max(1, count), np.int32)
        check(_L().simuse_pop_download_compact(self._h, _ptr(out, C.c_int32), int(count)))
        return out[:count]

    def download_spikes(self) -> np.ndarray:
        out = np.zeros(self.n, np.int32)
        check(_L().simuse_pop_download_spikes(self._h, _ptr(out, C.c_int32), out.size))
        return out

    def download_states(self) -> np.ndarray:
        out = np.zeros(self.ns * self.stride, np.float32)
        check(_L().simuse_pop_download_states(self._h, _ptr(out, C.c_float), out.size))
        return out

    def step(self, states: np.ndarray, params, i_in, spikes: np.ndarray, dt: float) -> None:
        """Naive step: upload, one step, states and spikes back in place."""
        self.run(states, params, i_in, spikes, dt, 1)

    def run(self, states: np.ndarray, params, i_in, spikes: np.ndarray, dt: float, n_steps: int) -> None:
        if states.dtype != np.float32 or states.size != self.ns * self.stride or not states.flags.c_contiguous:
            raise ValueError("states must be a contiguous float32 array of ns * stride values")
        if spikes.dtype != np.int32 or spikes.size != self.n or not spikes.flags.c_contiguous:
            raise ValueError("spikes must be a contiguous int32 array of n values")
        pr = _f32(params, self.np * self.stride, "params")
        cur = _f32(i_in, self.n, "i_in")
        check(_L().simuse_pop_run(self._h, _ptr(states, C.c_float), _ptr(pr, C.c_float), _ptr(cur, C.c_float),
                                  _ptr(spikes, C.c_int32), float(dt), int(n_steps)))
