"""Many independent networks stepped together on one device (simuse_batch_*).

Each network keeps the single-network step semantics (own ring, step
counter, RNG keys), so its results are bit-identical alone or in a batch.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import P, StdpParams, StpProfile, check, f32p, i32p, lib, u8p, u32p, u64p
from .network import N_PARAMS, N_STATE, NetworkPop, _descs, _ptr


class GenDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("n", C.c_uint32), ("k", C.c_uint32), ("seed", C.c_uint64),
                ("dt", C.c_float), ("exc_frac", C.c_float), ("w_exc", C.c_float), ("w_inh", C.c_float),
                ("d_min", C.c_uint8), ("d_max", C.c_uint8), ("drive_rate_hz", C.c_float),
                ("drive_weight", C.c_float), ("noise_sigma", C.c_float)]


_SIGS = {
    "simuse_batch_create": ([C.POINTER(P)], C.c_int),
    "simuse_batch_destroy": ([P], None),
    "simuse_batch_add_network": ([P, C.POINTER(_lib.PopDesc), C.c_int32, u32p, u32p, f32p, u8p, C.c_uint64, i32p],
                                 C.c_int),
    "simuse_batch_add_generated": ([P, C.POINTER(GenDesc), i32p], C.c_int),
    "simuse_batch_remove_network": ([P, C.c_int32], C.c_int),
    "simuse_batch_n_networks": ([P, i32p], C.c_int),
    "simuse_batch_add_poisson_source": ([P, C.c_int32, C.c_float, C.c_float, C.c_uint32, u32p, f32p, u8p,
                                         C.c_uint32], C.c_int),
    "simuse_batch_set_noise": ([P, C.c_int32, C.c_int32, f32p, C.c_uint64], C.c_int),
    "simuse_batch_set_plasticity": ([P, C.c_int32, C.POINTER(StdpParams), u8p], C.c_int),
    "simuse_batch_set_short_term_plasticity": ([P, C.c_int32, C.POINTER(StpProfile), C.c_int32, C.c_float, u8p,
                                                u8p, C.c_uint64], C.c_int),
    "simuse_batch_set_poisson_drive": ([P, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_uint32],
                                       C.c_int),
    "simuse_batch_set_horizon": ([P, C.c_int32, C.c_uint64], C.c_int),
    "simuse_batch_set_graph_steps": ([P, C.c_uint32], C.c_int),
    "simuse_batch_set_one_kernel": ([P, C.c_int32], C.c_int),
    "simuse_batch_get_one_kernel": ([P, i32p], C.c_int),
    "simuse_batch_set_scatter_mode": ([P, C.c_int32], C.c_int),
    "simuse_batch_autotune": ([P, C.c_uint32, i32p], C.c_int),
    "simuse_batch_set_block_tiles": ([P, C.c_uint32], C.c_int),
    "simuse_batch_step_async": ([P, C.c_uint32], C.c_int),
    "simuse_batch_sync": ([P], C.c_int),
    "simuse_batch_step_counter": ([P, C.c_int32, u64p], C.c_int),
    "simuse_batch_net_info": ([P, C.c_int32, u64p, u64p, i32p], C.c_int),
    "simuse_batch_pop_info": ([P, C.c_int32, C.c_int32, i32p, u64p], C.c_int),
    "simuse_batch_download_state": ([P, C.c_int32, C.c_int32, f32p, C.c_uint64], C.c_int),
    "simuse_batch_download_params": ([P, C.c_int32, C.c_int32, f32p, C.c_uint64], C.c_int),
    "simuse_batch_download_csr": ([P, C.c_int32, u32p, u32p, f32p, u8p], C.c_int),
    "simuse_batch_download_spike_ids": ([P, C.c_int32, u32p, C.c_uint32, u32p], C.c_int),
    "simuse_batch_counters": ([P, C.c_int32, u64p, u64p], C.c_int),
    "simuse_batch_record": ([P, C.c_int32, C.c_uint32], C.c_int),
    "simuse_batch_download_raster": ([P, C.c_int32, u64p, C.c_uint64, u32p, u64p], C.c_int),
    "simuse_batch_device_bytes": ([P, u64p], C.c_int),
    "simuse_batch_set_profiling": ([P, C.c_int32], C.c_int),
    "simuse_batch_kernel_times": ([P, f32p, C.c_int32], C.c_int),
    "simuse_gen_bytes": ([C.POINTER(GenDesc)], C.c_uint64),
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


def gen_desc(model, n, k, seed, dt, exc_frac=0.8, w_exc=0.5, w_inh=1.0, d_min=0, d_max=0,
             drive_rate_hz=0.0, drive_weight=0.0, noise_sigma=0.0) -> GenDesc:
    return GenDesc(int(model), int(n), int(k), int(seed) & ((1 << 64) - 1), float(dt), float(exc_frac),
                   float(w_exc), float(w_inh), int(d_min), int(d_max), float(drive_rate_hz), float(drive_weight),
                   float(noise_sigma))


def gen_bytes(g: GenDesc) -> int:
    return int(_L().simuse_gen_bytes(C.byref(g)))


class Batch:
    def __init__(self):
        h = C.c_void_p()
        check(_L().simuse_batch_create(C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None:
            _L().simuse_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_network(self, pops, row_ptr, col, w, delay) -> int:
        descs = _descs(pops)
        rp = np.ascontiguousarray(row_ptr, np.uint32)
        cc = np.ascontiguousarray(col, np.uint32)
        ww = np.ascontiguousarray(w, np.float32)
        dd = np.ascontiguousarray(delay, np.uint8)
        nid = C.c_int32()
        check(_L().simuse_batch_add_network(self._h, descs, len(pops), _ptr(rp, C.c_uint32), _ptr(cc, C.c_uint32),
                                            _ptr(ww, C.c_float), _ptr(dd, C.c_uint8), cc.size, C.byref(nid)))
        return int(nid.value)

    def add_spec(self, spec) -> int:
        nid = self.add_network(spec.pops, spec.row_ptr, spec.col, spec.w, spec.delay)
        for s in spec.sources:
            self.add_poisson_source(nid, s.rate_hz, s.dt_ms, s.seed, s.posts, s.weights, s.delays)
        for pop, sigma, seed in spec.noise:
            self.set_noise(nid, pop, sigma, seed)
        if spec.stdp:
            st = dict(spec.stdp)
            plastic = st.pop("plastic", None)
            sp = StdpParams(st["a_plus"], st["a_minus"], st["tau_plus"], st["tau_minus"], st["w_min"], st["w_max"])
            pl = None if plastic is None else np.ascontiguousarray(plastic, np.uint8)
            check(_L().simuse_batch_set_plasticity(self._h, nid, C.byref(sp),
                                                   _ptr(pl, C.c_uint8) if pl is not None else None))
        if spec.stp:
            profs = spec.stp["profiles"]
            arr = (StpProfile * len(profs))(*[StpProfile(*map(float, p)) for p in profs])
            syn = np.ascontiguousarray(spec.stp["syn_types"], np.uint8)
            src = spec.stp.get("src_types")
            src = None if src is None else np.ascontiguousarray(src, np.uint8)
            check(_L().simuse_batch_set_short_term_plasticity(
                self._h, nid, arr, len(profs), float(spec.stp["dt"]), _ptr(syn, C.c_uint8),
                _ptr(src, C.c_uint8) if src is not None else None, 0 if src is None else src.size))
        return nid

    def add_generated(self, g: GenDesc) -> int:
        nid = C.c_int32()
        check(_L().simuse_batch_add_generated(self._h, C.byref(g), C.byref(nid)))
        return int(nid.value)

    def remove(self, nid: int):
        check(_L().simuse_batch_remove_network(self._h, nid))

    def n_networks(self) -> int:
        n = C.c_int32()
        check(_L().simuse_batch_n_networks(self._h, C.byref(n)))
        return int(n.value)

    def add_poisson_source(self, nid, rate_hz, dt_ms, seed, posts, weights, delays):
        p = np.ascontiguousarray(posts, np.uint32)
        w = np.ascontiguousarray(weights, np.float32)
        d = np.ascontiguousarray(delays, np.uint8)
        check(_L().simuse_batch_add_poisson_source(self._h, nid, rate_hz, dt_ms, int(seed) & 0xFFFFFFFF,
                                                   _ptr(p, C.c_uint32), _ptr(w, C.c_float), _ptr(d, C.c_uint8),
                                                   p.size))

    def set_noise(self, nid, pop, sigma, seed):
        s = np.ascontiguousarray(sigma, np.float32)
        check(_L().simuse_batch_set_noise(self._h, nid, pop, _ptr(s, C.c_float), int(seed) & ((1 << 64) - 1)))

    def set_poisson_drive(self, nid, pop, rate_hz, dt_ms, weight, seed0):
        check(_L().simuse_batch_set_poisson_drive(self._h, nid, pop, rate_hz, dt_ms, weight,
                                                  int(seed0) & 0xFFFFFFFF))

    def set_horizon(self, nid, steps):
        check(_L().simuse_batch_set_horizon(self._h, nid, int(steps)))

    def set_graph_steps(self, steps):
        check(_L().simuse_batch_set_graph_steps(self._h, int(steps)))

    def set_scatter_mode(self, mode):
        check(_L().simuse_batch_set_scatter_mode(self._h, int(mode)))

    def set_block_tiles(self, tiles):
        check(_L().simuse_batch_set_block_tiles(self._h, int(tiles)))

    def set_one_kernel(self, on):
        check(_L().simuse_batch_set_one_kernel(self._h, 1 if on else 0))

    def one_kernel(self) -> bool:
        v = C.c_int32()
        check(_L().simuse_batch_get_one_kernel(self._h, C.byref(v)))
        return bool(v.value)

    def autotune(self, trial_steps=16) -> int:
        m = C.c_int32()
        check(_L().simuse_batch_autotune(self._h, int(trial_steps), C.byref(m)))
        return int(m.value)

    def step_async(self, n):
        check(_L().simuse_batch_step_async(self._h, int(n)))

    def sync(self):
        check(_L().simuse_batch_sync(self._h))

    def run(self, n):
        self.step_async(n)
        self.sync()

    def step_counter(self, nid) -> int:
        s = C.c_uint64()
        check(_L().simuse_batch_step_counter(self._h, nid, C.byref(s)))
        return int(s.value)

    def info(self, nid):
        n, ns, npops = C.c_uint64(), C.c_uint64(), C.c_int32()
        check(_L().simuse_batch_net_info(self._h, nid, C.byref(n), C.byref(ns), C.byref(npops)))
        return int(n.value), int(ns.value), int(npops.value)

    def pop_info(self, nid, pop):
        m, n = C.c_int32(), C.c_uint64()
        check(_L().simuse_batch_pop_info(self._h, nid, pop, C.byref(m), C.byref(n)))
        return int(m.value), int(n.value)

    def download_state(self, nid, pop):
        m, n = self.pop_info(nid, pop)
        s = (n + 7) // 8 * 8
        out = np.zeros((N_STATE[m], s), np.float32)
        check(_L().simuse_batch_download_state(self._h, nid, pop, _ptr(out, C.c_float), out.size))
        return out

    def download_params(self, nid, pop):
        m, n = self.pop_info(nid, pop)
        s = (n + 7) // 8 * 8
        out = np.zeros((N_PARAMS[m], s), np.float32)
        check(_L().simuse_batch_download_params(self._h, nid, pop, _ptr(out, C.c_float), out.size))
        return out

    def download_csr(self, nid):
        n, ns, _ = self.info(nid)
        rp = np.zeros(n + 1, np.uint32)
        col = np.zeros(max(ns, 1), np.uint32)
        w = np.zeros(max(ns, 1), np.float32)
        d = np.zeros(max(ns, 1), np.uint8)
        check(_L().simuse_batch_download_csr(self._h, nid, _ptr(rp, C.c_uint32), _ptr(col, C.c_uint32),
                                             _ptr(w, C.c_float), _ptr(d, C.c_uint8)))
        return rp, col[:ns], w[:ns], d[:ns]

    def pops(self, nid):
        """Current populations of a network as NetworkPop objects (for checkers)."""
        _, _, npops = self.info(nid)
        out = []
        for p in range(npops):
            m, n = self.pop_info(nid, p)
            out.append(NetworkPop(m, n, self.download_state(nid, p), self.download_params(nid, p)))
        return out

    def download_spike_ids(self, nid):
        cnt = C.c_uint32()
        check(_L().simuse_batch_download_spike_ids(self._h, nid, None, 0, C.byref(cnt)))
        out = np.zeros(max(1, cnt.value), np.uint32)
        check(_L().simuse_batch_download_spike_ids(self._h, nid, _ptr(out, C.c_uint32), out.size, C.byref(cnt)))
        return out[: cnt.value]

    def counters(self, nid):
        s, e = C.c_uint64(), C.c_uint64()
        check(_L().simuse_batch_counters(self._h, nid, C.byref(s), C.byref(e)))
        return int(s.value), int(e.value)

    def record(self, nid, max_steps):
        check(_L().simuse_batch_record(self._h, nid, int(max_steps)))

    def download_raster_bits(self, nid):
        steps, wps = C.c_uint32(), C.c_uint64()
        check(_L().simuse_batch_download_raster(self._h, nid, None, 0, C.byref(steps), C.byref(wps)))
        buf = np.zeros(max(1, steps.value * wps.value), np.uint64)
        check(_L().simuse_batch_download_raster(self._h, nid, _ptr(buf, C.c_uint64), buf.size, C.byref(steps),
                                                C.byref(wps)))
        return buf[: steps.value * wps.value].reshape(steps.value, wps.value)

    def download_raster(self, nid):
        """Dense int32 raster [steps, n_total] of a recording network (the bit
        raster holds per population ceil(n/256) tiles of four 64-bit words)."""
        bits = self.download_raster_bits(nid)
        T = bits.shape[0]
        n, _, npops = self.info(nid)
        out = np.zeros((T, n), np.int32)
        word0 = off = 0
        for p in range(npops):
            _, pn = self.pop_info(nid, p)
            nt = (pn + 255) // 256
            wp = np.ascontiguousarray(bits[:, word0: word0 + nt * 4])
            un = np.unpackbits(wp.view(np.uint8).reshape(T, -1), axis=1, bitorder="little")
            out[:, off: off + pn] = un[:, :pn]
            word0 += nt * 4
            off += pn
        return out

    def device_bytes(self) -> int:
        b = C.c_uint64()
        check(_L().simuse_batch_device_bytes(self._h, C.byref(b)))
        return int(b.value)

    def set_profiling(self, on: bool):
        check(_L().simuse_batch_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        out = (C.c_float * 4)()
        check(_L().simuse_batch_kernel_times(self._h, out, 4))
        return [float(x) for x in out]
