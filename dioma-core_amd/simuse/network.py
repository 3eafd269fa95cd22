"""Host-side mirror of the reference's GPU network API.

``CompiledNetwork`` has the method names and argument meaning of
``pillars_cuda::CompiledNetwork`` (crates/SiMuse/crates/pillars-cuda/src/
network.rs:182-882): ``new`` (the constructor), ``step``, ``step_graph``,
``sync``, ``download_spikes``, ``download_states``, ``upload_weights``,
``download_weights``, ``inject_current``, ``step_counter``, ``reset``,
``n_total``, ``n_pops``, ``pop_n``.  Where the reference returns
``Err(String)`` this raises :class:`SimuseError` with the library's message.
Steps follow the CPU canon ``pillars::api::Network::step``
(crates/SiMuse/crates/pillars/src/api/net/network/step.rs:18-189).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import PopDesc, StdpParams, StpProfile, check, lib

N_STATE = {_lib.MODEL_LIF: 2, _lib.MODEL_IZHIKEVICH: 3, _lib.MODEL_HH: 6, _lib.MODEL_GPEA: 9, _lib.MODEL_ALPHA_LIF: 3}
N_PARAMS = {_lib.MODEL_LIF: 6, _lib.MODEL_IZHIKEVICH: 7, _lib.MODEL_HH: 13, _lib.MODEL_GPEA: 14,
            _lib.MODEL_ALPHA_LIF: 7}


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def stride_for(n: int) -> int:
    """Column stride of the reference's SoA layout: ceil8(n)."""
    return (n + 7) // 8 * 8


@dataclass
class NetworkPop:
    """One population (mirrors ``NetworkPop``, PC/src/network.rs:142-156).

    ``states``/``params`` are column-major SoA arrays of shape
    ``(ns, stride)`` / ``(np, stride)``.
    """

    model: int
    n: int
    states: np.ndarray
    params: np.ndarray

    @property
    def ns(self) -> int:
        return N_STATE[self.model]

    @property
    def np_(self) -> int:
        return N_PARAMS[self.model]

    @property
    def stride(self) -> int:
        return int(self.states.shape[1])

    def desc(self) -> PopDesc:
        st = np.ascontiguousarray(self.states, dtype=np.float32)
        pr = np.ascontiguousarray(self.params, dtype=np.float32)
        if st.shape != (self.ns, self.stride) or pr.shape != (self.np_, self.stride):
            raise ValueError("state/param arrays must be (ns, stride) and (np, stride)")
        self._keep = (st, pr)
        return PopDesc(self.model, self.n, self.ns, self.np_, self.stride, _ptr(st, C.c_float), _ptr(pr, C.c_float))


def _descs(pops: Sequence[NetworkPop]):
    arr = (PopDesc * len(pops))()
    for i, p in enumerate(pops):
        arr[i] = p.desc()
    return arr


class CompiledNetwork:
    """A spiking network resident on one MI355X (one HIP stream per handle)."""

    def __init__(self, pops: Sequence[NetworkPop], row_ptr, col_idx, weights, delays, dt: float = 0.0):
        self._h = None
        self.pops = list(pops)
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        ci = np.ascontiguousarray(col_idx, dtype=np.uint32)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        d = np.ascontiguousarray(delays, dtype=np.uint8)
        if not (len(ci) == len(w) == len(d)):
            raise ValueError("col_idx/weights/delays length mismatch")
        self.n_syn = int(len(ci))
        descs = _descs(self.pops)
        h = C.c_void_p()
        check(lib().simuse_net_create(descs, len(self.pops), _ptr(rp, C.c_uint32), _ptr(ci, C.c_uint32),
                                      _ptr(w, C.c_float), _ptr(d, C.c_uint8), self.n_syn, float(dt), C.byref(h)))
        self._h = h
        self._n_total = sum(p.n for p in self.pops)

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if self._h is not None:
            lib().simuse_net_destroy(self._h)
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

    # -- configuration (before the first step) --------------------------------
    def add_poisson_source(self, rate_hz: float, dt_ms: float, seed: int, posts, weights, delays) -> None:
        p = np.ascontiguousarray(posts, dtype=np.uint32)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        d = np.ascontiguousarray(delays, dtype=np.uint8)
        check(lib().simuse_net_add_poisson_source(self._h, float(rate_hz), float(dt_ms), int(seed) & 0xFFFFFFFF,
                                                  _ptr(p, C.c_uint32), _ptr(w, C.c_float), _ptr(d, C.c_uint8), len(p)))

    def set_noise(self, pop: int, sigma, seed: int) -> None:
        s = np.ascontiguousarray(sigma, dtype=np.float32)
        if len(s) != self.pops[pop].n:
            raise ValueError("sigma must have one entry per neuron")
        check(lib().simuse_net_set_noise(self._h, pop, _ptr(s, C.c_float), int(seed) & ((1 << 64) - 1)))

    def set_poisson_drive(self, pop: int, rate_hz: float, dt_ms: float, weight: float, seed0: int) -> None:
        """One Poisson source per neuron (seed seed0 + i*2654435761 mod 2^32, one
        edge of `weight`, delay <= 1), rolled inside the neuron kernel."""
        check(lib().simuse_net_set_poisson_drive(self._h, pop, float(rate_hz), float(dt_ms), float(weight),
                                                 int(seed0) & 0xFFFFFFFF))

    def set_plasticity(self, a_plus, a_minus, tau_plus, tau_minus, w_min, w_max, plastic=None) -> None:
        sp = StdpParams(a_plus, a_minus, tau_plus, tau_minus, w_min, w_max)
        if plastic is None:
            check(lib().simuse_net_set_plasticity(self._h, C.byref(sp), None))
        else:
            m = np.ascontiguousarray(plastic, dtype=np.uint8)
            check(lib().simuse_net_set_plasticity(self._h, C.byref(sp), _ptr(m, C.c_uint8)))

    def set_short_term_plasticity(self, profiles, dt: float, syn_types, src_types=None) -> None:
        """Tsodyks-Markram synapses: profiles = [(u0, tau_rec, tau_fac or 0)], per-edge profile
        ids (1-based, 0 = static) for the synapses and optionally the source edges
        (Network::register_plasticity, core.rs:597-619; tm_spike, push.rs:60-86)."""
        arr = (StpProfile * len(profiles))(*[StpProfile(*map(float, p)) for p in profiles])
        syn = np.ascontiguousarray(syn_types, dtype=np.uint8)
        if syn.size != self.n_syn:
            raise ValueError("syn_types must have one entry per synapse")
        src = None if src_types is None else np.ascontiguousarray(src_types, dtype=np.uint8)
        check(lib().simuse_net_set_short_term_plasticity(
            self._h, arr, len(profiles), float(dt), _ptr(syn, C.c_uint8),
            _ptr(src, C.c_uint8) if src is not None else None, 0 if src is None else src.size))

    def set_scatter_mode(self, mode: int) -> None:
        check(lib().simuse_net_set_scatter_mode(self._h, int(mode)))

    def scatter_mode(self) -> int:
        m = C.c_int32()
        check(lib().simuse_net_get_scatter_mode(self._h, C.byref(m)))
        return int(m.value)

    def set_block_tiles(self, tiles: int) -> None:
        check(lib().simuse_net_set_block_tiles(self._h, int(tiles)))

    def block_tiles(self) -> int:
        g = C.c_uint32()
        check(lib().simuse_net_get_block_tiles(self._h, C.byref(g)))
        return int(g.value)

    def autotune(self, trial_steps: int = 16) -> int:
        check(lib().simuse_net_autotune(self._h, int(trial_steps)))
        return self.scatter_mode()

    def set_graph_steps(self, steps: int) -> None:
        check(lib().simuse_net_set_graph_steps(self._h, int(steps)))

    def set_one_kernel(self, on: bool) -> None:
        """One-kernel step (True, the default) or the binned form for networks that can take either."""
        check(lib().simuse_net_set_one_kernel(self._h, 1 if on else 0))

    # -- stepping ---------------------------------------------------------------
    def step(self) -> None:
        """One step, synchronous (CompiledNetwork::step, network.rs:411)."""
        check(lib().simuse_net_step(self._h))

    def stdp_step(self, trace_decay: float, lr_ltp: float, lr_ltd: float, w_min: float, w_max: float) -> None:
        """One step with the dense every-synapse STDP rule (CompiledNetwork::stdp_step,
        network.rs:495-718); asynchronous."""
        check(lib().simuse_net_stdp_step(self._h, float(trace_decay), float(lr_ltp), float(lr_ltd),
                                         float(w_min), float(w_max)))

    def step_graph(self, n_steps: int = 1) -> None:
        """Enqueue n steps of graph replay and return (network.rs:431)."""
        check(lib().simuse_net_step_async(self._h, int(n_steps)))

    def run(self, n_steps: int) -> None:
        self.step_graph(n_steps)
        self.sync()

    def sync(self) -> None:
        check(lib().simuse_net_sync(self._h))

    # -- transfers ----------------------------------------------------------------
    def download_spikes(self) -> np.ndarray:
        out = np.zeros(self._n_total, dtype=np.int32)
        check(lib().simuse_net_download_spikes(self._h, _ptr(out, C.c_int32), out.size))
        return out

    def download_spike_ids(self) -> np.ndarray:
        cnt = C.c_uint32()
        check(lib().simuse_net_download_spike_ids(self._h, None, 0, C.byref(cnt)))
        out = np.zeros(max(1, cnt.value), dtype=np.uint32)
        check(lib().simuse_net_download_spike_ids(self._h, _ptr(out, C.c_uint32), out.size, C.byref(cnt)))
        return out[: cnt.value]

    def download_states(self, pop: int) -> np.ndarray:
        p = self.pops[pop]
        out = np.zeros((p.ns, p.stride), dtype=np.float32)
        check(lib().simuse_net_download_state(self._h, pop, _ptr(out, C.c_float), out.size))
        return out

    def download_ring(self) -> np.ndarray:
        out = np.zeros((32, self._n_total), dtype=np.float32)
        check(lib().simuse_net_download_ring(self._h, _ptr(out, C.c_float), out.size))
        return out

    def upload_weights(self, weights) -> None:
        w = np.ascontiguousarray(weights, dtype=np.float32)
        check(lib().simuse_net_upload_weights(self._h, _ptr(w, C.c_float), w.size))

    def download_weights(self) -> np.ndarray:
        out = np.zeros(self.n_syn, dtype=np.float32)
        check(lib().simuse_net_download_weights(self._h, _ptr(out, C.c_float), out.size))
        return out

    def inject_current(self, pop: int, indices, values) -> None:
        i = np.ascontiguousarray(indices, dtype=np.uint32)
        v = np.ascontiguousarray(values, dtype=np.float32)
        if i.size != v.size:
            raise ValueError("inject_current: indices/values length mismatch")
        check(lib().simuse_net_inject_current(self._h, pop, _ptr(i, C.c_uint32), _ptr(v, C.c_float), i.size))

    # -- electrical coupling (before the first step) ------------------------------
    def connect_gap_junction(self, from_pop: int, from_idx: int, to_pop: int, to_idx: int,
                             conductance: float) -> None:
        """Network::connect_gap_junction (crates/SiMuse/crates/pillars/src/api/net/network/core.rs:391)."""
        check(lib().simuse_net_connect_gap_junction(self._h, from_pop, from_idx, to_pop, to_idx, float(conductance)))

    def connect_graded_synapse(self, from_pop: int, from_idx: int, to_pop: int, to_idx: int,
                               conductance: float) -> None:
        """Network::connect_graded_synapse (core.rs:438)."""
        check(lib().simuse_net_connect_graded_synapse(self._h, from_pop, from_idx, to_pop, to_idx,
                                                      float(conductance)))

    def add_gap_junctions(self, a, b, g) -> None:
        """Gap junctions between global neurons a[k], b[k] with conductance g[k], in order."""
        a, b, g = (np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32),
                   np.ascontiguousarray(g, np.float32))
        check(lib().simuse_net_add_gap_junctions(self._h, _ptr(a, C.c_uint32), _ptr(b, C.c_uint32),
                                                 _ptr(g, C.c_float), a.size))

    def add_graded_synapses(self, pre, post, g) -> None:
        pre, post, g = (np.ascontiguousarray(pre, np.uint32), np.ascontiguousarray(post, np.uint32),
                        np.ascontiguousarray(g, np.float32))
        check(lib().simuse_net_add_graded_synapses(self._h, _ptr(pre, C.c_uint32), _ptr(post, C.c_uint32),
                                                   _ptr(g, C.c_float), pre.size))

    def register_sigmoid_graded(self, vth: float, delta: float, k: float, erev: float, conductance: float,
                                dt_ms: float) -> int:
        """Network::register_sigmoid_graded (core.rs:483); returns the profile id."""
        out = C.c_int32()
        check(lib().simuse_net_register_sigmoid_graded(self._h, vth, delta, k, erev, conductance, dt_ms,
                                                       C.byref(out)))
        return int(out.value)

    def connect_sigmoid_graded(self, profile: int, from_pop: int, from_idx: int, to_pop: int, to_idx: int,
                               weight: float) -> None:
        check(lib().simuse_net_connect_sigmoid_graded(self._h, profile, from_pop, from_idx, to_pop, to_idx,
                                                      float(weight)))

    def add_sigmoid_graded(self, profile: int, pre, post, weight) -> None:
        pre, post, w = (np.ascontiguousarray(pre, np.uint32), np.ascontiguousarray(post, np.uint32),
                        np.ascontiguousarray(weight, np.float32))
        check(lib().simuse_net_add_sigmoid_graded(self._h, profile, _ptr(pre, C.c_uint32), _ptr(post, C.c_uint32),
                                                  _ptr(w, C.c_float), pre.size))

    def inject_voltage(self, pop: int, indices, values) -> None:
        """WgpuNetwork::inject_voltage (crates/SiMuse/crates/pillars-wgpu/src/network.rs:1632-1659):
        v += value on local neurons of ``pop`` before the next step; ``indices`` None
        means one value per neuron of the population."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        if indices is None:
            check(lib().simuse_net_inject_voltage(self._h, pop, None, _ptr(v, C.c_float), v.size))
            return
        i = np.ascontiguousarray(indices, dtype=np.uint32)
        if i.size != v.size:
            raise ValueError("inject_voltage: indices/values length mismatch")
        check(lib().simuse_net_inject_voltage(self._h, pop, _ptr(i, C.c_uint32), _ptr(v, C.c_float), i.size))

    def reset(self, pops: Sequence[NetworkPop]) -> None:
        self.pops = list(pops)
        descs = _descs(self.pops)
        check(lib().simuse_net_reset(self._h, descs, len(self.pops)))

    # -- receptor synapses ----------------------------------------------------------
    def set_receptors(self, types, scale: float = 1.0, ext=None) -> None:
        """Conductance-based synapses: receptor type per synapse (0 AMPA, 1 NMDA,
        2 GABA_A, 3 GABA_B, 4 nAChR, 5 GluCl, 6 mAChR, 255 current-based), jump
        scale, optional constant external current per neuron."""
        t = np.ascontiguousarray(types, dtype=np.uint8)
        if t.size != self.n_syn:
            raise ValueError(f"receptor types: {t.size} values for {self.n_syn} synapses")
        x = None
        if ext is not None:
            x = np.ascontiguousarray(ext, dtype=np.float32)
            if x.size != self._n_total:
                raise ValueError("external current: one value per neuron")
        check(lib().simuse_net_set_receptors(self._h, _ptr(t, C.c_uint8), float(scale),
                                             _ptr(x, C.c_float) if x is not None else None))
        self._rec_keep = (t, x)

    def download_conductances(self) -> np.ndarray:
        out = np.zeros((7, self._n_total), dtype=np.float32)
        check(lib().simuse_net_download_conductances(self._h, _ptr(out, C.c_float), out.size))
        return out

    # -- checkpoint / resume --------------------------------------------------------
    def save_state(self, path) -> None:
        """Write the network's whole dynamic state to a ``.taas`` v2 file."""
        check(lib().simuse_net_save_state(self._h, os.fsencode(path)))

    def load_state(self, path) -> None:
        """Restore a state written by :meth:`save_state` into a network built
        the same way; stepping then continues exactly as the saved one."""
        check(lib().simuse_net_load_state(self._h, os.fsencode(path)))

    # -- introspection -------------------------------------------------------------
    def step_counter(self) -> int:
        s = C.c_uint64()
        check(lib().simuse_net_step_counter(self._h, C.byref(s)))
        return int(s.value)

    def n_total(self) -> int:
        return self._n_total

    def n_pops(self) -> int:
        return len(self.pops)

    def pop_n(self, pop: int) -> int:
        return self.pops[pop].n if 0 <= pop < len(self.pops) else 0

    def counters(self):
        s, e = C.c_uint64(), C.c_uint64()
        check(lib().simuse_net_counters(self._h, C.byref(s), C.byref(e)))
        return int(s.value), int(e.value)

    def device_bytes(self) -> int:
        b = C.c_uint64()
        check(lib().simuse_net_device_bytes(self._h, C.byref(b)))
        return int(b.value)

    def step_form(self) -> int:
        """FORM_ONE_KERNEL, FORM_SPLIT, FORM_BINNED or FORM_CHAINED (finalizes the network)."""
        f = C.c_int32()
        check(lib().simuse_net_step_form(self._h, C.byref(f)))
        return int(f.value)

    # -- recording and statistics ----------------------------------------------
    def record(self, max_steps: int) -> None:
        check(lib().simuse_net_record(self._h, int(max_steps)))

    def download_raster(self) -> np.ndarray:
        """Dense int32 raster [steps, n_total] from the device bit raster."""
        words = C.c_uint64()
        check(lib().simuse_net_raster_words(self._h, C.byref(words)))
        steps = C.c_uint32()
        check(lib().simuse_net_download_raster(self._h, None, 0, C.byref(steps)))
        buf = np.zeros(max(1, steps.value * words.value), dtype=np.uint64)
        check(lib().simuse_net_download_raster(self._h, _ptr(buf, C.c_uint64), buf.size, C.byref(steps)))
        T, W = int(steps.value), int(words.value)
        bits = buf[: T * W].reshape(T, W)
        out = np.zeros((T, self._n_total), dtype=np.int32)
        word0 = 0
        off = 0
        for p in self.pops:
            nt = (p.n + 255) // 256
            wp = bits[:, word0: word0 + nt * 4]
            un = np.unpackbits(wp.view(np.uint8).reshape(T, -1), axis=1, bitorder="little")
            out[:, off: off + p.n] = un[:, : p.n]
            word0 += nt * 4
            off += p.n
        return out

    def raster_stats(self, first: int, count: int, dt_s: float):
        rates = np.zeros(count, dtype=np.float32)
        cvs = np.zeros(count, dtype=np.float32)
        check(lib().simuse_net_raster_stats(self._h, first, count, float(dt_s), _ptr(rates, C.c_float),
                                            _ptr(cvs, C.c_float)))
        return rates, cvs

    def raster_stats_noalloc(self, first: int, count: int, dt_s: float):
        """raster_stats through the entry point that allocates no device memory
        (at most RASTER_STATS_NOALLOC_MAX neurons per call)."""
        rates = np.zeros(count, dtype=np.float32)
        cvs = np.zeros(count, dtype=np.float32)
        check(lib().simuse_net_raster_stats_noalloc(self._h, first, count, float(dt_s), _ptr(rates, C.c_float),
                                                    _ptr(cvs, C.c_float)))
        return rates, cvs

    # -- profiling -------------------------------------------------------------------
    def set_profiling(self, on: bool) -> None:
        check(lib().simuse_net_set_profiling(self._h, 1 if on else 0))

    def time_steps(self, n: int) -> float:
        """Run n steps (captured graphs) between two HIP events on the
        network's stream; returns the elapsed milliseconds."""
        out = C.c_float()
        check(lib().simuse_net_time_steps(self._h, int(n), C.byref(out)))
        return float(out.value)

    def time_steps_queued(self, n: int) -> float:
        """time_steps with the steps queued behind a device-side spacer: the
        events bracket the step kernels only (no host submission gap)."""
        out = C.c_float()
        check(lib().simuse_net_time_steps_queued(self._h, int(n), C.byref(out)))
        return float(out.value)

    def kernel_times(self):
        """Mean ms per step of [neuron (+ fused scatter), stdp, insert, fix-up] over profiled steps."""
        out = (C.c_float * 4)()
        check(lib().simuse_net_kernel_times(self._h, out, 4))
        return [float(x) for x in out]
