"""Population builders with the reference's SoA column order.

Columns follow the ``#[state]``/``#[param]`` field order of each model:
  LIF        P/src/zoo/neurons/lif.rs:22-40
             state [v, refr]; params [dt, inv_tau_m, v_rest, v_thresh, v_reset, tau_ref]
  Izhikevich P/src/zoo/neurons/izhikevich.rs:42-64
             state [v, u, refr]; params [dt, a, b, c, d, tau_ref, v_thresh]
  HH         field order of P/src/zoo/neurons/hh.rs:12-32 (the CPU stub) with
             the equations of PC/kernels/hh.cu:107-216
             state [v, m, h, n, refr, c]; params [dt, g_na, g_k, g_leak, e_na, e_k,
             e_leak, inv_cm, q10, v_thresh, g_ca_t, e_ca_t, tau_ref]
  alpha-LIF  P/src/zoo/neurons/alpha_lif.rs:60-91
             state [v, g, refr_cnt]; params [dt, time_factor, v_rest, v_thresh,
             v_reset, refr_steps, syn_decay]
(crates/SiMuse/crates/ is the prefix of every path above.)
"""
from __future__ import annotations

import math
from typing import Sequence, Union

import numpy as np

from ._lib import MODEL_ALPHA_LIF, MODEL_GPEA, MODEL_HH, MODEL_IZHIKEVICH, MODEL_LIF
from .network import N_PARAMS, N_STATE, NetworkPop, stride_for

Num = Union[float, np.ndarray]


def _cols(n: int, values: Sequence[Num]) -> np.ndarray:
    s = stride_for(n)
    out = np.zeros((len(values), s), dtype=np.float32)
    for k, v in enumerate(values):
        out[k, :n] = np.asarray(v, dtype=np.float32) if np.ndim(v) else np.float32(v)
    return out


def make_pop(model: int, n: int, params: Sequence[Num], states: Sequence[Num]) -> NetworkPop:
    if len(params) != N_PARAMS[model] or len(states) != N_STATE[model]:
        raise ValueError("wrong number of columns for the model")
    return NetworkPop(model, n, _cols(n, states), _cols(n, params))


def lif_pop(n: int, dt: Num, inv_tau_m: Num, v_rest: Num, v_thresh: Num, v_reset: Num, tau_ref: Num,
            v0: Num, refr0: Num = 0.0) -> NetworkPop:
    return make_pop(MODEL_LIF, n, [dt, inv_tau_m, v_rest, v_thresh, v_reset, tau_ref], [v0, refr0])


def izh_pop(n: int, dt: Num, a: Num, b: Num, c: Num, d: Num, tau_ref: Num, v_thresh: Num, v0: Num,
            u0: Num, refr0: Num = 0.0) -> NetworkPop:
    return make_pop(MODEL_IZHIKEVICH, n, [dt, a, b, c, d, tau_ref, v_thresh], [v0, u0, refr0])


def round_half_away(x: np.float32) -> np.float32:
    """f32::round: halves away from zero (numpy rounds them to even).  An FP32
    value plus 0.5 is exact in FP64."""
    x64 = float(x)
    r = math.floor(x64 + 0.5) if x64 >= 0 else -math.floor(-x64 + 0.5)
    return np.float32(r)


def alpha_lif_pop(n: int, dt: float = 0.5, tau_ref: float = 2.2, tau_m: float = 20.0, v_rest: float = -52.0,
                  v_thresh: float = -45.0, v_reset: float = -52.0, tau_syn: float = 3.0) -> NetworkPop:
    """Alpha-LIF population with the derived parameters and initial state of
    ensure_built (crates/dioma-engine/src/brain/dynamics/gpu_pool.rs:205-240):
    refr_steps = max(round(tau_ref/dt), 1), time_factor = dt/tau_m,
    syn_decay = 1 - dt/tau_syn, all in FP32; v = v_rest, g = 0, refr_cnt = refr_steps."""
    f = np.float32
    refr = max(round_half_away(f(tau_ref) / f(dt)), f(1.0))
    tf = f(dt) / f(tau_m)
    syn_decay = f(1.0) - f(dt) / f(tau_syn)
    return make_pop(MODEL_ALPHA_LIF, n, [dt, tf, v_rest, v_thresh, v_reset, refr, syn_decay], [v_rest, 0.0, refr])


# NEURON reference geometry of the GPEA tests (soma L=20, diam=40 um): area pi*40*20 um^2
GPEA_AREA = 2513.2742
# 0.5 nA IClamp over that area, in uA/cm^2
GPEA_PULSE = 0.5e-3 / (GPEA_AREA * 1e-8)


def gpea_pop(n: int, dt: Num = 0.025, celsius: Num = 24.0, inv_cm: Num = 1.0, v_thresh: Num = -20.0,
             area: Num = GPEA_AREA, gnabar: Num = 49.0, gkdrbar: Num = 57.0, gl: Num = 0.35, el: Num = -60.0,
             gcatbar: Num = 5.0, gkcabar: Num = 1.0, kca: Num = 2.0, vol: Num = 3.355e-11, ca_gain: Num = 0.1,
             v0: Num = -60.0, m0: Num = 0.0414, h0: Num = 0.906, n0: Num = 0.2165, p0: Num = 0.355,
             q0: Num = 0.01325, cai0: Num = 5e-5, r0: Num = 0.1821261, was_below0: Num = 1.0) -> NetworkPop:
    """GPEA (pGPeA_k.mod) population with the NEURON reference parameters and initial
    state of crates/SiMuse/crates/pillars/src/zoo/neurons/gpea.rs:363-424 (V=-60 mV, gates at
    steady state, cai = 5e-5 as NEURON's finitialize sets it, dt 0.025 ms, 24 C)."""
    return make_pop(MODEL_GPEA, n,
                    [dt, celsius, inv_cm, v_thresh, area, gnabar, gkdrbar, gl, el, gcatbar, gkcabar, kca, vol,
                     ca_gain],
                    [v0, m0, h0, n0, p0, q0, cai0, r0, was_below0])


def hh_pop(n: int, dt: Num, g_na: Num = 120.0, g_k: Num = 36.0, g_leak: Num = 0.3, e_na: Num = 50.0,
           e_k: Num = -77.0, e_leak: Num = -65.0, inv_cm: Num = 1.0, q10: Num = 1.0, v_thresh: Num = 0.0,
           g_ca_t: Num = 0.0, e_ca_t: Num = 120.0, tau_ref: Num = 2.0, v0: Num = -65.0, m0: Num = 0.0529,
           h0: Num = 0.596, n0: Num = 0.318, c0: Num = 0.0) -> NetworkPop:
    """Standard HH (parameters and resting gates of
    crates/SiMuse/crates/pillars-connectome/src/sim/hh.rs:21-38)."""
    return make_pop(MODEL_HH, n,
                    [dt, g_na, g_k, g_leak, e_na, e_k, e_leak, inv_cm, q10, v_thresh, g_ca_t, e_ca_t, tau_ref],
                    [v0, m0, h0, n0, 0.0, c0])
