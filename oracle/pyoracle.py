"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker and the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcPop(C.Structure):
    _fields_ = [("model", C.c_int32), ("n", C.c_uint32), ("ns", C.c_uint32), ("np", C.c_uint32),
                ("stride", C.c_uint32), ("states", C.POINTER(C.c_float)), ("params", C.POINTER(C.c_float))]


class OrcStdp(C.Structure):
    _fields_ = [("a_plus", C.c_float), ("a_minus", C.c_float), ("tau_plus", C.c_float),
                ("tau_minus", C.c_float), ("w_min", C.c_float), ("w_max", C.c_float)]


_lib = None
P = C.c_void_p
u32p, u64p, i32p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_int32)
f32p, f64p, u8p = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_uint8)

_SIGS = {
    "orc_net_create": ([C.c_int, C.POINTER(OrcPop), C.c_int32, u32p, u32p, f32p, u8p, C.c_uint64], P),
    "orc_net_destroy": ([P], None),
    "orc_net_add_poisson_source": ([P, C.c_float, C.c_float, C.c_uint32, u32p, f32p, u8p, C.c_uint32], C.c_int),
    "orc_net_set_noise": ([P, C.c_int32, f32p, C.c_uint64], C.c_int),
    "orc_net_set_plasticity": ([P, C.POINTER(OrcStdp), u8p], C.c_int),
    "orc_net_inject_current": ([P, C.c_int32, u32p, f32p, C.c_uint32], C.c_int),
    "orc_net_inject_voltage": ([P, C.c_int32, u32p, f32p, C.c_uint32], C.c_int),
    "orc_net_add_gap_junction": ([P, C.c_uint32, C.c_uint32, C.c_float], None),
    "orc_net_add_graded_synapse": ([P, C.c_uint32, C.c_uint32, C.c_float], None),
    "orc_net_add_sigmoid_profile": ([P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float], C.c_int),
    "orc_net_add_sigmoid_conn": ([P, C.c_int32, C.c_uint32, C.c_uint32, C.c_float], C.c_int),
    "orc_pool_create": ([C.c_void_p, C.c_int32], P),
    "orc_pool_destroy": ([P], None),
    "orc_pool_register": ([P, C.c_uint32, i32p, i32p, f32p, f32p], C.c_int32),
    "orc_pool_set_weights": ([P, C.c_int32, f32p, C.c_uint64], C.c_int),
    "orc_pool_step": ([P, f32p, C.c_uint32, u32p], C.c_int),
    "orc_pool_set_delays": ([P, C.c_int32, u8p, C.c_uint64], C.c_int),
    "orc_pool_state": ([P, C.c_int32, f32p], C.c_int),
    "orc_pool_n_total": ([P], C.c_uint32),
    "orc_pool_delay": ([P], C.c_int32),
    "orc_net_step": ([P, C.c_uint32], None),
    "orc_net_stdp_step": ([P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float], None),
    "orc_net_step_counter": ([P], C.c_uint64),
    "orc_net_spikes": ([P, i32p], None),
    "orc_net_active": ([P, u32p], C.c_uint32),
    "orc_net_state": ([P, C.c_int32, C.c_int32, f64p], C.c_int),
    "orc_net_ring": ([P, f64p], None),
    "orc_net_weights": ([P, f64p], None),
    "orc_net_total_spikes": ([P], C.c_uint64),
    "orc_net_total_events": ([P], C.c_uint64),
    "orc_net_record": ([P, C.c_int32], None),
    "orc_net_raster_steps": ([P], C.c_uint64),
    "orc_net_raster": ([P, i32p], None),
    "orc_net_set_threads": ([P, C.c_int32], None),
    "orc_poisson_spike": ([C.c_uint32, C.c_uint32, C.c_uint64, C.c_float], C.c_int32),
    "orc_gauss": ([C.c_uint64, C.c_uint32, C.c_uint64], C.c_float),
    "orc_tm_spike": ([C.c_float, C.c_float, C.c_float, C.c_int32, f32p, f32p, u32p, C.c_float, C.c_uint64],
                     C.c_float),
    "orc_powi_f32": ([C.c_float, C.c_int32], C.c_float),
    "orc_fast_exp": ([C.c_float], C.c_float),
    "orc_gpea_rates": ([C.c_float, C.c_float, f32p], None),
    "orc_net_set_stp": ([P, C.c_void_p, C.c_int32, C.c_float, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int),
    "orc_cos2pi_f32": ([C.c_float], C.c_float),
    "orc_sin2pi_f32": ([C.c_float], C.c_float),
    "orc_exp_f32": ([C.c_float], C.c_float),
    "orc_ln_f32": ([C.c_float], C.c_float),
    "orc_vtrap_f32": ([C.c_float, C.c_float], C.c_float),
    "orc_vtrap_f64": ([C.c_double, C.c_double], C.c_double),
    "orc_compact": ([i32p, C.c_uint32, u32p], C.c_uint32),
    "orc_net_set_receptors": ([P, C.c_void_p, C.c_float, C.c_void_p], C.c_int),
    "orc_net_conductances": ([P, f64p], None),
    "orc_net_set_state": ([P, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_void_p], C.c_int),
    "orc_rec_decay_factors": ([C.c_float, f32p], None),
    "orc_nmda_block": ([C.c_float], C.c_float),
    "orc_rec_currents": ([f32p, f32p, f32p, C.c_uint32], None),
    "orc_kura_create": ([C.c_int, C.c_uint32, u32p, u32p, f32p, f32p, f32p, C.c_float], P),
    "orc_kura_destroy": ([P], None),
    "orc_kura_step_rk4": ([P, C.c_uint32], None),
    "orc_kura_step_euler": ([P, C.c_uint32], None),
    "orc_kura_phases": ([P, f64p], None),
    "orc_kura_order": ([P], C.c_double),
    "orc_firing_rates": ([i32p, C.c_uint32, C.c_uint32, C.c_float, f32p], None),
    "orc_isi_cv": ([i32p, C.c_uint32, C.c_uint32, C.c_float, f32p], None),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class OrcStp(C.Structure):
    _fields_ = [("u0", C.c_float), ("tau_rec", C.c_float), ("tau_fac", C.c_float)]


def stp_arrays(spec):
    """Profiles and per-edge type arrays of spec.stp = {profiles: [(u0, tau_rec, tau_fac)],
    dt, syn_types (n_syn), src_types (all source edges, optional)}."""
    st = spec.stp
    profs = st["profiles"]
    arr = (OrcStp * len(profs))(*[OrcStp(*map(float, p)) for p in profs])
    syn = np.ascontiguousarray(st["syn_types"], np.uint8)
    src = st.get("src_types")
    n_src = sum(len(s.posts) for s in spec.sources)
    if src is not None:
        src = np.ascontiguousarray(src, np.uint8)
    return arr, len(profs), syn, src, n_src


def tm_spike(u0, a_rec, a_fac, is_fac, state, weight, step):
    """tm_spike on one synapse; state = [r, u, last] is advanced in place."""
    r, u, last = C.c_float(state[0]), C.c_float(state[1]), C.c_uint32(state[2])
    out = lib().orc_tm_spike(u0, a_rec, a_fac, 1 if is_fac else 0, C.byref(r), C.byref(u), C.byref(last),
                             weight, step)
    state[0], state[1], state[2] = r.value, u.value, last.value
    return float(out)


def powi_f32(a, n):
    return float(lib().orc_powi_f32(a, n))


def fast_exp(x):
    return float(lib().orc_fast_exp(x))


def sin2pi_f32(u):
    return float(lib().orc_sin2pi_f32(u))


def rec_decay_factors(dt):
    out = np.zeros(7, np.float32)
    lib().orc_rec_decay_factors(dt, _p(out, C.c_float))
    return out


def nmda_block(v):
    return float(lib().orc_nmda_block(v))


def rec_currents(g, v, i_out):
    """SynapticKinetics::compute_currents: i_out += currents of g[7][n] at v."""
    g = np.ascontiguousarray(g, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.array(i_out, dtype=np.float32, copy=True)
    lib().orc_rec_currents(_p(g, C.c_float), _p(v, C.c_float), _p(out, C.c_float), v.size)
    return out


def gpea_rates(v, cai):
    """m_inf, h_inf, n_inf, p_inf, q_inf, r_inf, tau_m, tau_h, tau_n, tau_p, tau_q."""
    out = np.zeros(11, np.float32)
    lib().orc_gpea_rates(v, cai, _p(out, C.c_float))
    return out


class OracleNetwork:
    """CPU restatement of Network::step built from a NetSpec (kind 32 or 64)."""

    def __init__(self, spec, kind: int = 32, threads: int = 1):
        L = lib()
        self.spec = spec
        self.kind = kind
        arr = (OrcPop * len(spec.pops))()
        keep = []
        for i, p in enumerate(spec.pops):
            st = np.ascontiguousarray(p.states, dtype=np.float32)
            pr = np.ascontiguousarray(p.params, dtype=np.float32)
            keep += [st, pr]
            arr[i] = OrcPop(p.model, p.n, st.shape[0], pr.shape[0], st.shape[1], _p(st, C.c_float),
                            _p(pr, C.c_float))
        rp = np.ascontiguousarray(spec.row_ptr, np.uint32)
        col = np.ascontiguousarray(spec.col, np.uint32)
        w = np.ascontiguousarray(spec.w, np.float32)
        d = np.ascontiguousarray(spec.delay, np.uint8)
        self.h = L.orc_net_create(kind, arr, len(spec.pops), _p(rp, C.c_uint32), _p(col, C.c_uint32),
                                  _p(w, C.c_float), _p(d, C.c_uint8), col.size)
        if not self.h:
            raise ValueError("oracle rejected the network")
        self.N = spec.n_total
        for s in spec.sources:
            pp = np.ascontiguousarray(s.posts, np.uint32)
            ww = np.ascontiguousarray(s.weights, np.float32)
            dd = np.ascontiguousarray(s.delays, np.uint8)
            rc = L.orc_net_add_poisson_source(self.h, s.rate_hz, s.dt_ms, s.seed & 0xFFFFFFFF, _p(pp, C.c_uint32),
                                              _p(ww, C.c_float), _p(dd, C.c_uint8), pp.size)
            assert rc == 0
        for pop, sigma, seed in spec.noise:
            sg = np.ascontiguousarray(sigma, np.float32)
            assert L.orc_net_set_noise(self.h, pop, _p(sg, C.c_float), seed) == 0
        if spec.stdp:
            st = dict(spec.stdp)
            plastic = st.pop("plastic", None)
            if plastic is None:
                plastic = np.ones(col.size, np.uint8)
            plastic = np.ascontiguousarray(plastic, np.uint8)
            sp = OrcStdp(st["a_plus"], st["a_minus"], st["tau_plus"], st["tau_minus"], st["w_min"], st["w_max"])
            assert L.orc_net_set_plasticity(self.h, C.byref(sp), _p(plastic, C.c_uint8)) == 0
        if getattr(spec, "stp", None):
            prof, n_prof, syn, src, n_src = stp_arrays(spec)
            assert L.orc_net_set_stp(self.h, prof, n_prof, float(spec.stp["dt"]), syn.ctypes.data,
                                     src.ctypes.data if src is not None else None, n_src) == 0
            self._stp_keep = (prof, syn, src)
        rec = getattr(spec, "receptors", None)
        if rec:
            types = np.ascontiguousarray(rec["types"], np.uint8)
            ext = rec.get("ext")
            ext = np.ascontiguousarray(ext, np.float32) if ext is not None else None
            assert L.orc_net_set_receptors(self.h, types.ctypes.data, float(rec.get("scale", 1.0)),
                                           ext.ctypes.data if ext is not None else None) == 0
        cp = getattr(spec, "coupling", None)
        if cp:
            if cp.get("gap") is not None:
                for a, b, g in zip(*cp["gap"]):
                    L.orc_net_add_gap_junction(self.h, int(a), int(b), float(g))
            if cp.get("graded") is not None:
                for a, b, g in zip(*cp["graded"]):
                    L.orc_net_add_graded_synapse(self.h, int(a), int(b), float(g))
            for prof, (pre, post, w) in cp.get("sigmoid", []):
                pid = L.orc_net_add_sigmoid_profile(self.h, *[float(x) for x in prof])
                for a, b, x in zip(pre, post, w):
                    assert L.orc_net_add_sigmoid_conn(self.h, pid, int(a), int(b), float(x)) == 0
        if threads > 1:
            L.orc_net_set_threads(self.h, threads)

    def close(self):
        if self.h:
            lib().orc_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        lib().orc_net_step(self.h, n)

    def stdp_step(self, decay, lr_ltp, lr_ltd, w_min, w_max):
        """One step with the dense every-synapse STDP (CompiledNetwork::stdp_step)."""
        lib().orc_net_stdp_step(self.h, decay, lr_ltp, lr_ltd, w_min, w_max)

    def inject_current(self, pop, idx, val):
        i = np.ascontiguousarray(idx, np.uint32)
        v = np.ascontiguousarray(val, np.float32)
        assert lib().orc_net_inject_current(self.h, pop, _p(i, C.c_uint32), _p(v, C.c_float), i.size) == 0

    def inject_voltage(self, pop, idx, val):
        """v += val (idx None: one value per neuron of pop) before the next step."""
        v = np.ascontiguousarray(val, np.float32)
        if idx is None:
            assert lib().orc_net_inject_voltage(self.h, pop, None, _p(v, C.c_float), v.size) == 0
            return
        i = np.ascontiguousarray(idx, np.uint32)
        assert lib().orc_net_inject_voltage(self.h, pop, _p(i, C.c_uint32), _p(v, C.c_float), i.size) == 0

    def step_counter(self):
        return int(lib().orc_net_step_counter(self.h))

    def spikes(self):
        out = np.zeros(self.N, np.int32)
        lib().orc_net_spikes(self.h, _p(out, C.c_int32))
        return out

    def active(self):
        out = np.zeros(self.N + 1, np.uint32)
        k = lib().orc_net_active(self.h, _p(out, C.c_uint32))
        return out[:k].copy()

    def state(self, pop, col):
        out = np.zeros(self.spec.pops[pop].n, np.float64)
        assert lib().orc_net_state(self.h, pop, col, _p(out, C.c_double)) == 0
        return out

    def ring(self):
        out = np.zeros((32, self.N), np.float64)
        lib().orc_net_ring(self.h, _p(out, C.c_double))
        return out

    def weights(self):
        out = np.zeros(self.spec.n_syn, np.float64)
        lib().orc_net_weights(self.h, _p(out, C.c_double))
        return out

    def set_state_from_snapshot(self, path):
        """Load a device state written by CompiledNetwork.save_state (`.taas`
        v2, every FP32 value exact in either twin)."""
        from simuse.formats import TaasReader
        r = TaasReader.from_file(path)
        keep = []

        def ptr(a):
            if a is None:
                return None
            keep.append(a)
            return a.ctypes.data

        pops = [np.ascontiguousarray(r.get_f32(f"pop{p}.states")) for p in range(len(self.spec.pops))]
        arr = (C.c_void_p * len(pops))(*[ptr(x) for x in pops])
        w64 = r.get_u64("stdp.w64").view(np.float64) if r.has_section("stdp.w64") else None
        pre = r.get_f32("stdp.pre") if r.has_section("stdp.pre") else None
        post = r.get_f32("stdp.post") if r.has_section("stdp.post") else None
        g = r.get_f32("receptors.g") if r.has_section("receptors.g") else None
        rc = lib().orc_net_set_state(self.h, int(r.get_u64("step")[0]), C.cast(arr, C.c_void_p),
                                     ptr(r.get_f32("ring")), ptr(r.get_f32("weights")), ptr(w64), ptr(pre),
                                     ptr(post), ptr(g))
        assert rc == 0

    def conductances(self):
        out = np.zeros((7, self.N), np.float64)
        lib().orc_net_conductances(self.h, _p(out, C.c_double))
        return out

    def counters(self):
        return int(lib().orc_net_total_spikes(self.h)), int(lib().orc_net_total_events(self.h))

    def record(self, on=True):
        lib().orc_net_record(self.h, 1 if on else 0)

    def raster(self):
        t = int(lib().orc_net_raster_steps(self.h))
        out = np.zeros((t, self.N), np.int32)
        if t:
            lib().orc_net_raster(self.h, _p(out, C.c_int32))
        return out


class OracleKuramoto:
    def __init__(self, row_ptr, col, w, omega, theta0, dt, kind=64):
        rp = np.ascontiguousarray(row_ptr, np.uint32)
        cc = np.ascontiguousarray(col, np.uint32)
        ww = np.ascontiguousarray(w, np.float32)
        om = np.ascontiguousarray(omega, np.float32)
        th = np.ascontiguousarray(theta0, np.float32)
        self.n = om.size
        self.h = lib().orc_kura_create(kind, self.n, _p(rp, C.c_uint32), _p(cc, C.c_uint32), _p(ww, C.c_float),
                                       _p(om, C.c_float), _p(th, C.c_float), dt)

    def rk4(self, n):
        lib().orc_kura_step_rk4(self.h, n)

    def euler(self, n):
        lib().orc_kura_step_euler(self.h, n)

    def phases(self):
        out = np.zeros(self.n, np.float64)
        lib().orc_kura_phases(self.h, _p(out, C.c_double))
        return out

    def order(self):
        return float(lib().orc_kura_order(self.h))

    def __del__(self):
        try:
            lib().orc_kura_destroy(self.h)
        except Exception:
            pass


def poisson_spike(seed, stream, step, p):
    return bool(lib().orc_poisson_spike(seed, stream, step, p))


def gauss(seed, i, step):
    return float(lib().orc_gauss(seed, i, step))


def cos2pi_f32(u):
    return float(lib().orc_cos2pi_f32(u))


def exp_f32(x):
    return float(lib().orc_exp_f32(x))


def ln_f32(x):
    return float(lib().orc_ln_f32(x))


def vtrap_f32(x, y):
    return float(lib().orc_vtrap_f32(x, y))


def vtrap_f64(x, y):
    return float(lib().orc_vtrap_f64(x, y))


def compact(spikes):
    s = np.ascontiguousarray(spikes, np.int32)
    out = np.zeros(s.size + 1, np.uint32)
    k = lib().orc_compact(_p(s, C.c_int32), s.size, _p(out, C.c_uint32))
    return out[:k].copy()


def firing_rates(raster, dt_s):
    r = np.ascontiguousarray(raster, np.int32)
    t, n = r.shape
    out = np.zeros(n, np.float32)
    lib().orc_firing_rates(_p(r, C.c_int32), n, t, dt_s, _p(out, C.c_float))
    return out


def isi_cv(raster, dt_s):
    r = np.ascontiguousarray(raster, np.int32)
    t, n = r.shape
    out = np.zeros(n, np.float32)
    lib().orc_isi_cv(_p(r, C.c_int32), n, t, dt_s, _p(out, C.c_float))
    return out


class _OrcPoolParams(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("dt_ms", "tau_ref", "tau_m", "v_rest", "v_thresh", "v_reset", "tau_syn",
                                          "t_delay")]


class OraclePool:
    """The creature pool (gpu_pool.rs + csr_scatter.wgsl) restated on the CPU:
    same register / set_weights / step / state calls as the HIP pool, with the
    tick input given per creature."""

    DEFAULT = dict(dt_ms=0.5, tau_ref=2.2, tau_m=20.0, v_rest=-52.0, v_thresh=-45.0, v_reset=-52.0, tau_syn=3.0,
                   t_delay=1.8)

    def __init__(self, config=None, delay_steps=-1):
        cfg = dict(self.DEFAULT)
        if config:
            cfg.update(config)
        p = _OrcPoolParams(**{k: float(v) for k, v in cfg.items()})
        self.h = lib().orc_pool_create(C.byref(p), int(delay_steps))
        self.slots = []

    def close(self):
        if self.h:
            lib().orc_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def register(self, row_ptr, col, values, nt):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        c = np.ascontiguousarray(col, np.int32)
        w = np.ascontiguousarray(values, np.float32)
        t = np.ascontiguousarray(nt, np.float32)
        n = rp.size - 1
        k = lib().orc_pool_register(self.h, n, _p(rp, C.c_int32), _p(c, C.c_int32), _p(w, C.c_float),
                                    _p(t, C.c_float))
        off = sum(s[1] for s in self.slots)
        soff = sum(s[3] for s in self.slots)
        self.slots.append((off, n, soff, int(rp[-1])))
        return k

    @property
    def n(self):
        return int(lib().orc_pool_n_total(self.h))

    @property
    def delay(self):
        return int(lib().orc_pool_delay(self.h))

    def set_weights(self, k, w):
        a = np.ascontiguousarray(w, np.float32)
        assert lib().orc_pool_set_weights(self.h, int(k), _p(a, C.c_float), a.size) == 0

    def set_delays(self, k, d):
        a = np.ascontiguousarray(d, np.uint8)
        assert lib().orc_pool_set_delays(self.h, int(k), _p(a, C.c_uint8), a.size) == 0

    def step(self, exts, spt):
        n = self.n
        ext = np.zeros(max(n, 1), np.float32)
        for (off, m, _, _), e in zip(self.slots, exts):
            if e is not None:
                ext[off: off + m] = np.asarray(e, np.float32)
        out = np.zeros(max(1, spt * n), np.uint32)
        assert lib().orc_pool_step(self.h, _p(ext, C.c_float), int(spt), _p(out, C.c_uint32)) == 0
        return out[: spt * n]

    def state(self, col):
        out = np.zeros(max(self.n, 1), np.float32)
        assert lib().orc_pool_state(self.h, int(col), _p(out, C.c_float)) == 0
        return out[: self.n]
