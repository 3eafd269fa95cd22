"""ctypes binding of libsimuse.so (the C ABI declared in include/simuse.h).

Loading fails loudly: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libsimuse.so")

OK = 0
E_INVALID = -1
E_HIP = -2
E_NOMEM = -3
E_STATE = -4
E_UNSUPPORTED = -5
E_DEVICE = -6

MODEL_LIF = 0
MODEL_IZHIKEVICH = 1
MODEL_HH = 2
MODEL_GPEA = 3
MODEL_ALPHA_LIF = 4

POOL_MAX_STEPS_PER_SUBMIT = 64
RASTER_STATS_NOALLOC_MAX = 65536

KURA_EULER = 0
KURA_RK4 = 1

FORM_ONE_KERNEL = 0
FORM_BINNED = 1
FORM_CHAINED = 2
FORM_SPLIT = 3

SCATTER_AUTO = 0
SCATTER_TILE = 1
SCATTER_WAVE = 2


class PopDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("n", C.c_uint32),
        ("ns", C.c_uint32),
        ("np", C.c_uint32),
        ("stride", C.c_uint32),
        ("states", C.POINTER(C.c_float)),
        ("params", C.POINTER(C.c_float)),
    ]


class StdpParams(C.Structure):
    _fields_ = [
        ("a_plus", C.c_float),
        ("a_minus", C.c_float),
        ("tau_plus", C.c_float),
        ("tau_minus", C.c_float),
        ("w_min", C.c_float),
        ("w_max", C.c_float),
    ]


class StpProfile(C.Structure):
    _fields_ = [("u0", C.c_float), ("tau_rec", C.c_float), ("tau_fac", C.c_float)]


class PoolConfig(C.Structure):
    """NeuronConfig::LifAlpha (crates/SiMuse/crates/pillars-connectome/src/runner.rs:114)."""
    _fields_ = [(k, C.c_float) for k in ("dt_ms", "tau_ref", "tau_m", "v_rest", "v_thresh", "v_reset", "tau_syn",
                                          "t_delay")]


class PoolSlot(C.Structure):
    """PoolSlot (crates/dioma-engine/src/brain/dynamics/gpu_pool.rs:5-11)."""
    _fields_ = [("neuron_off", C.c_uint64), ("n", C.c_uint64), ("syn_off", C.c_uint64), ("n_syn", C.c_uint64)]


class SimuseError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[simuse {code}] {msg}")
        self.code = code


_lib = None

P = C.c_void_p
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)

_SIGS = {
    "simuse_last_error": ([], C.c_char_p),
    "simuse_device_count": ([i32p], C.c_int),
    "simuse_net_create": ([C.POINTER(PopDesc), C.c_int32, u32p, u32p, f32p, u8p, C.c_uint64, C.c_float,
                           C.POINTER(P)], C.c_int),
    "simuse_net_destroy": ([P], None),
    "simuse_net_add_poisson_source": ([P, C.c_float, C.c_float, C.c_uint32, u32p, f32p, u8p, C.c_uint32], C.c_int),
    "simuse_net_set_noise": ([P, C.c_int32, f32p, C.c_uint64], C.c_int),
    "simuse_net_set_poisson_drive": ([P, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_uint32], C.c_int),
    "simuse_net_set_plasticity": ([P, C.POINTER(StdpParams), u8p], C.c_int),
    "simuse_net_set_scatter_mode": ([P, C.c_int32], C.c_int),
    "simuse_net_get_scatter_mode": ([P, i32p], C.c_int),
    "simuse_net_autotune": ([P, C.c_uint32], C.c_int),
    "simuse_net_stdp_step": ([P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float], C.c_int),
    "simuse_net_set_short_term_plasticity": ([P, C.POINTER(StpProfile), C.c_int32, C.c_float, u8p, u8p,
                                              C.c_uint64], C.c_int),
    "simuse_net_set_block_tiles": ([P, C.c_uint32], C.c_int),
    "simuse_net_get_block_tiles": ([P, u32p], C.c_int),
    "simuse_net_set_graph_steps": ([P, C.c_uint32], C.c_int),
    "simuse_net_set_one_kernel": ([P, C.c_int32], C.c_int),
    "simuse_net_step": ([P], C.c_int),
    "simuse_net_step_async": ([P, C.c_uint32], C.c_int),
    "simuse_net_sync": ([P], C.c_int),
    "simuse_net_download_spikes": ([P, i32p, C.c_uint64], C.c_int),
    "simuse_net_download_spike_ids": ([P, u32p, C.c_uint32, u32p], C.c_int),
    "simuse_net_download_state": ([P, C.c_int32, f32p, C.c_uint64], C.c_int),
    "simuse_net_download_ring": ([P, f32p, C.c_uint64], C.c_int),
    "simuse_net_upload_weights": ([P, f32p, C.c_uint64], C.c_int),
    "simuse_net_download_weights": ([P, f32p, C.c_uint64], C.c_int),
    "simuse_net_inject_current": ([P, C.c_int32, u32p, f32p, C.c_uint32], C.c_int),
    "simuse_net_inject_voltage": ([P, C.c_int32, u32p, f32p, C.c_uint32], C.c_int),
    "simuse_net_connect_gap_junction": ([P, C.c_int32, C.c_uint32, C.c_int32, C.c_uint32, C.c_float], C.c_int),
    "simuse_net_add_gap_junctions": ([P, u32p, u32p, f32p, C.c_uint32], C.c_int),
    "simuse_net_connect_graded_synapse": ([P, C.c_int32, C.c_uint32, C.c_int32, C.c_uint32, C.c_float], C.c_int),
    "simuse_net_add_graded_synapses": ([P, u32p, u32p, f32p, C.c_uint32], C.c_int),
    "simuse_net_register_sigmoid_graded": ([P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                            i32p], C.c_int),
    "simuse_net_connect_sigmoid_graded": ([P, C.c_int32, C.c_int32, C.c_uint32, C.c_int32, C.c_uint32, C.c_float],
                                          C.c_int),
    "simuse_net_add_sigmoid_graded": ([P, C.c_int32, u32p, u32p, f32p, C.c_uint32], C.c_int),
    "simuse_net_register_kinetics": ([P, C.c_int32, f32p, C.c_int32, C.c_float, i32p], C.c_int),
    "simuse_pool_create": ([C.POINTER(PoolConfig), C.c_int32, C.POINTER(P)], C.c_int),
    "simuse_pool_destroy": ([P], None),
    "simuse_pool_register": ([P, C.c_uint32, i32p, i32p, f32p, f32p, C.POINTER(PoolSlot)], C.c_int),
    "simuse_pool_set_delays": ([P, C.c_uint32, u8p, C.c_uint64], C.c_int),
    "simuse_pool_n_creatures": ([P, u32p], C.c_int),
    "simuse_pool_slots": ([P, C.POINTER(PoolSlot), C.c_uint32], C.c_int),
    "simuse_pool_sync_weights": ([P, C.POINTER(f32p), u8p, C.c_uint32], C.c_int),
    "simuse_pool_step": ([P, C.POINTER(f32p), C.c_uint32, C.c_uint32, u32p, C.c_uint64], C.c_int),
    "simuse_pool_download_state": ([P, C.c_int32, f32p, C.c_uint64], C.c_int),
    "simuse_pool_info": ([P, u64p, u64p, u64p, i32p], C.c_int),
    "simuse_pool_time_substeps": ([P, C.c_uint32, f32p], C.c_int),
    "simuse_net_reset": ([P, C.POINTER(PopDesc), C.c_int32], C.c_int),
    "simuse_net_set_receptors": ([P, u8p, C.c_float, f32p], C.c_int),
    "simuse_net_download_conductances": ([P, f32p, C.c_uint64], C.c_int),
    "simuse_net_save_state": ([P, C.c_char_p], C.c_int),
    "simuse_net_load_state": ([P, C.c_char_p], C.c_int),
    "simuse_net_step_counter": ([P, u64p], C.c_int),
    "simuse_net_n_total": ([P, u64p], C.c_int),
    "simuse_net_n_pops": ([P, i32p], C.c_int),
    "simuse_net_pop_n": ([P, C.c_int32, u64p], C.c_int),
    "simuse_net_counters": ([P, u64p, u64p], C.c_int),
    "simuse_net_record": ([P, C.c_uint32], C.c_int),
    "simuse_net_raster_words": ([P, u64p], C.c_int),
    "simuse_net_download_raster": ([P, u64p, C.c_uint64, u32p], C.c_int),
    "simuse_net_raster_stats": ([P, C.c_uint32, C.c_uint32, C.c_float, f32p, f32p], C.c_int),
    "simuse_net_raster_stats_noalloc": ([P, C.c_uint32, C.c_uint32, C.c_float, f32p, f32p], C.c_int),
    "simuse_net_device_bytes": ([P, u64p], C.c_int),
    "simuse_net_step_form": ([P, i32p], C.c_int),
    "simuse_net_set_profiling": ([P, C.c_int32], C.c_int),
    "simuse_net_kernel_times": ([P, f32p, C.c_int32], C.c_int),
    "simuse_net_time_steps": ([P, C.c_uint32, f32p], C.c_int),
    "simuse_net_time_steps_queued": ([P, C.c_uint32, f32p], C.c_int),
}

# Optional families (present when their sources are built in).
_SIGS_OPT = {
    "simuse_kura_create": ([u32p, u32p, f32p, f32p, f32p, C.c_uint32, C.c_float, C.c_int32, C.POINTER(P)], C.c_int),
    "simuse_kura_destroy": ([P], None),
    "simuse_kura_set_group": ([P, C.c_int32], C.c_int),
    "simuse_kura_step_async": ([P, C.c_uint32], C.c_int),
    "simuse_kura_sync": ([P], C.c_int),
    "simuse_kura_download_phases": ([P, f64p, C.c_uint32], C.c_int),
    "simuse_kura_order_parameter": ([P, f64p], C.c_int),
    "simuse_kura_phase_coherence": ([P, f32p, f64p], C.c_int),
    "simuse_kura_step_counter": ([P, u64p], C.c_int),
    "simuse_kura_set_profiling": ([P, C.c_int32], C.c_int),
    "simuse_kura_kernel_times": ([P, f32p, C.c_int32], C.c_int),
}


def lib():
    """Load libsimuse.so once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python dioma-core_amd/build_lib.py` "
            "(the HIP library is required; there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    for name, (args, res) in _SIGS_OPT.items():
        f = getattr(L, name, None)
        if f is not None:
            f.argtypes = args
            f.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().simuse_last_error()
        raise SimuseError(rc, msg.decode() if msg else "")


def device_count() -> int:
    n = C.c_int32(0)
    check(lib().simuse_device_count(C.byref(n)))
    return int(n.value)
