"""Creature pool: the game engine's BrainPool on the HIP pool kernels.

Mirrors crates/dioma-engine/src/brain/dynamics/gpu_pool.rs (BrainPool): many
small alpha-LIF brains registered one by one, stepped together as one
block-diagonal network with the per-post pull scatter of
crates/SiMuse/crates/pillars-wgpu/shaders/csr_scatter.wgsl:26-59.  ``step``
returns the spike history of the tick, ``spt`` rows of ``n_neurons`` u32,
and the creature slots, like the reference's ``(&[u32], &[PoolSlot])``.
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple, Optional, Sequence, Tuple

import numpy as np

from ._lib import PoolConfig, PoolSlot, check, lib


class Slot(NamedTuple):
    """PoolSlot (gpu_pool.rs:5-11)."""
    neuron_off: int
    n: int
    syn_off: int
    n_syn: int


# pool_config (gpu_pool.rs:13-24)
POOL_CONFIG = dict(dt_ms=0.5, tau_ref=2.2, tau_m=20.0, v_rest=-52.0, v_thresh=-45.0, v_reset=-52.0, tau_syn=3.0,
                   t_delay=1.8)


def _f32(a, n: Optional[int] = None, what: str = "array") -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and x.size != n:
        raise ValueError(f"{what}: {x.size} values, expected {n}")
    return x


class BrainPool:
    """BrainPool::new / register / sync_weights / step (gpu_pool.rs:57-203)."""

    def __init__(self, config: Optional[dict] = None, delay_steps: int = -1):
        cfg = dict(POOL_CONFIG)
        if config:
            cfg.update(config)
        self.config = cfg
        c = PoolConfig(**{k: float(v) for k, v in cfg.items()})
        h = C.c_void_p()
        check(lib().simuse_pool_create(C.byref(c), int(delay_steps), C.byref(h)))
        self._h = h
        self._slots: list = []

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().simuse_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- creatures ----------------------------------------------------------------
    def register(self, row_ptr, col_indices, values, nt_types) -> Slot:
        """Append a creature given its ConnectomeData CSR (local indices)."""
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        n = rp.size - 1
        if n < 0:
            raise ValueError("register: row_ptr needs n + 1 entries")
        ci = np.ascontiguousarray(col_indices, dtype=np.int32)
        w = _f32(values)
        nt = _f32(nt_types, n, "nt_types")
        if ci.size != w.size or ci.size != int(rp[-1]):
            raise ValueError("register: col_indices/values must hold row_ptr[n] entries")
        s = PoolSlot()
        check(lib().simuse_pool_register(self._h, n, rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                         ci.ctypes.data_as(C.POINTER(C.c_int32)),
                                         w.ctypes.data_as(C.POINTER(C.c_float)),
                                         nt.ctypes.data_as(C.POINTER(C.c_float)), C.byref(s)))
        slot = Slot(int(s.neuron_off), int(s.n), int(s.syn_off), int(s.n_syn))
        self._slots.append(slot)
        return slot

    def set_delays(self, creature: int, delays) -> None:
        d = np.ascontiguousarray(delays, dtype=np.uint8)
        check(lib().simuse_pool_set_delays(self._h, int(creature), d.ctypes.data_as(C.POINTER(C.c_uint8)), d.size))

    def n_creatures(self) -> int:
        return len(self._slots)

    def slots(self) -> list:
        return list(self._slots)

    @property
    def n_neurons(self) -> int:
        return sum(s.n for s in self._slots)

    def info(self) -> Tuple[int, int, int, int]:
        """(neurons, synapses, substeps since the last rebuild, delay)."""
        n, s, t = C.c_uint64(), C.c_uint64(), C.c_uint64()
        d = C.c_int32()
        check(lib().simuse_pool_info(self._h, C.byref(n), C.byref(s), C.byref(t), C.byref(d)))
        return int(n.value), int(s.value), int(t.value), int(d.value)

    # -- ticks --------------------------------------------------------------------
    def sync_weights(self, mirrors: Sequence[Tuple[Optional[np.ndarray], bool]]) -> None:
        """mirrors[k] = (weights of creature k, dirty)."""
        if len(mirrors) != len(self._slots):
            raise ValueError("sync_weights: one mirror per creature")
        keep = []
        ptrs = (C.POINTER(C.c_float) * max(1, len(mirrors)))()
        dirty = np.zeros(max(1, len(mirrors)), dtype=np.uint8)
        for k, ((m, d), s) in enumerate(zip(mirrors, self._slots)):
            if d:
                a = _f32(m, s.n_syn, f"mirror {k}")
                keep.append(a)
                ptrs[k] = a.ctypes.data_as(C.POINTER(C.c_float))
                dirty[k] = 1
        check(lib().simuse_pool_sync_weights(self._h, ptrs, dirty.ctypes.data_as(C.POINTER(C.c_uint8)),
                                             len(mirrors)))

    def step(self, exts: Sequence[Optional[np.ndarray]], spt: int,
             readback: bool = True) -> Tuple[np.ndarray, list]:
        """One tick of ``spt`` substeps; returns (flat spikes [spt * n_neurons] u32, slots)."""
        if not self._slots:
            return np.zeros(0, dtype=np.uint32), []
        if len(exts) != len(self._slots):
            raise ValueError("step: one input (or None) per creature")
        keep = []
        ptrs = (C.POINTER(C.c_float) * len(exts))()
        for k, (e, s) in enumerate(zip(exts, self._slots)):
            if e is not None:
                a = _f32(e, s.n, f"ext {k}")
                keep.append(a)
                ptrs[k] = a.ctypes.data_as(C.POINTER(C.c_float))
        n = self.n_neurons
        flat = np.zeros(int(spt) * n if readback else 0, dtype=np.uint32)
        fp = flat.ctypes.data_as(C.POINTER(C.c_uint32)) if readback else None
        check(lib().simuse_pool_step(self._h, ptrs, len(exts), int(spt), fp, flat.size))
        return flat, self.slots()

    def time_substeps(self, n: int) -> float:
        """n more substeps (submits of <= 64, the last tick's input at each
        submit's start, no readback) between two HIP events; milliseconds."""
        out = C.c_float()
        check(lib().simuse_pool_time_substeps(self._h, int(n), C.byref(out)))
        return float(out.value)

    def download_state(self, col: int) -> np.ndarray:
        """State column of every neuron: 0 v, 1 g, 2 refr_cnt."""
        out = np.zeros(self.n_neurons, dtype=np.float32)
        check(lib().simuse_pool_download_state(self._h, int(col), out.ctypes.data_as(C.POINTER(C.c_float)),
                                               out.size))
        return out
