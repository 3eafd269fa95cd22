"""The oracle's conductance-based receptor synapses, pinned (CPU only).

Restates crates/SiMuse/crates/pillars/src/kernels/synapse/tests.rs:3-112
(decay factors, exponential decay, ohmic direction, NMDA Mg2+ block profile,
GABA_B / mAChR driving force, 8-lane vs scalar-tail consistency), then pins
what those leave open with closed forms: exp(-dt/tau) decay over many steps,
the Mg2+ block against 1/(1+0.33 e^{-0.06 v}), the exact step at which a
spike's conductance jump acts (CN/src/sim/hh.rs:115-166: jump after the
update, decay before the next current), the receptor code mapping, the
jump scale, and the lane/tail association split at n & ~7.
"""
import numpy as np
import pytest

import pyoracle as po
from simuse.models import lif_pop
from simuse.workloads import NetSpec, csr_from_triplets

TAU = np.array([2.0, 50.0, 10.0, 200.0, 3.0, 5.0, 150.0])
EREV = np.array([0.0, 0.0, -65.0, -90.0, 0.0, -70.0, -90.0], np.float32)


def currents(g, v, base=None):
    n = len(v)
    return po.rec_currents(g, v, np.zeros(n, np.float32) if base is None else base)


def test_decay_factors_precision():
    f = po.rec_decay_factors(0.5)
    assert abs(f[0] - 0.7788) < 0.001
    # every receptor against exp(-dt/tau) (fast_exp: relative error < 2e-5)
    for dt in (0.01, 0.025, 0.1, 0.5, 1.0, 2.0):
        f = po.rec_decay_factors(dt).astype(np.float64)
        assert np.all(np.abs(f / np.exp(-dt / TAU) - 1) < 2e-5), dt


def test_decay_all_exponential():
    f = po.rec_decay_factors(2.0)
    assert abs(np.float32(1.0) * f[0] - 0.3679) < 0.001
    assert abs(np.float32(2.0) * f[2] - 1.6375) < 0.001


def test_compute_currents_ohmic_direction():
    g = np.zeros((7, 2), np.float32)
    g[0, 0] = 0.1  # AMPA at -65: +6.5 (depolarising)
    g[2, 1] = 0.2  # GABA_A at rest: shunt, 0
    i = currents(g, np.array([-65.0, -65.0], np.float32))
    assert abs(i[0] - 6.5) < 0.01 and abs(i[1]) < 0.01


def test_nmda_mg_block_voltage_profile():
    v = np.array([-80.0, -65.0, -40.0, -20.0, 0.0, 20.0, 40.0], np.float32)
    g = np.zeros((7, 7), np.float32)
    g[1] = 1.0
    i = currents(g, v)
    assert i[0] < 5.0 and abs(i[4]) < 1.0 and i[5] < -5.0
    # closed form, and the clamp of the exponent at +-10
    for x in np.linspace(-100, 100, 401).astype(np.float32):
        b = 1.0 / (1.0 + 0.33 * np.exp(-0.06 * float(x)))
        assert abs(po.nmda_block(float(x)) - b) <= 2e-5 * b + 1e-7
    assert po.nmda_block(-400.0) == po.nmda_block(-10.0 / 0.06 - 1.0)
    assert po.nmda_block(400.0) == po.nmda_block(10.0 / 0.06 + 1.0)


def test_gabab_and_machr_inhibitory():
    g = np.zeros((7, 2), np.float32)
    g[3, 0] = 1.0
    g[6, 1] = 1.0
    i = currents(g, np.array([-65.0, -65.0], np.float32))
    assert abs(i[0] + 25.0) < 0.1 and abs(i[1] + 25.0) < 0.1


def test_simd_tail_consistent():
    n = 13  # 8 lanes + 5 tail
    g = np.zeros((7, n), np.float32)
    k = np.arange(n, dtype=np.float32) + 1
    g[0] = 0.1 * k
    g[1] = 0.05 * k
    g[2] = 0.02
    g[4] = 0.01
    v = (-70.0 + np.arange(n)).astype(np.float32)
    i = currents(g, v)
    for idx in (0, 4, 12):
        b = po.nmda_block(float(v[idx]))
        exp = (0.1 * k[idx] * (0 - v[idx]) + 0.05 * k[idx] * b * (0 - v[idx]) + 0.02 * (-65 - v[idx])
               + 0.01 * (0 - v[idx]))
        assert abs(i[idx] - exp) < 1e-4


def _orders(g, v, base):
    """Both association orders of compute_currents, in float32 step by step."""
    f32 = np.float32
    lane, tail = [], []
    for i in range(len(v)):
        b = f32(po.nmda_block(float(v[i])))
        t = [f32(f32(g[r, i] * b) if r == 1 else g[r, i]) * f32(EREV[r] - v[i]) for r in range(7)]
        cur = f32(base[i])
        for x in t:
            cur = f32(cur + f32(x))
        acc = f32(0)
        for x in t:
            acc = f32(acc + f32(x))
        lane.append(cur)
        tail.append(f32(f32(base[i]) + acc))
    return np.array(lane, np.float32), np.array(tail, np.float32)


@pytest.mark.parametrize("n", [8, 13, 21, 5])
def test_lane_tail_association_split(n):
    # kinetics.rs:164-203: below n & ~7 the input takes each receptor current
    # in turn, above it the sum of the seven; pick values where they differ
    rng = np.random.default_rng(n)
    g = (rng.random((7, n)) * 3).astype(np.float32)
    v = rng.uniform(-80, 30, n).astype(np.float32)
    base = rng.uniform(-1e3, 1e3, n).astype(np.float32)
    out = po.rec_currents(g, v, base)
    lane, tail = _orders(g, v, base)
    n8 = n & ~7
    assert np.array_equal(out[:n8].view(np.uint32), lane[:n8].view(np.uint32))
    assert np.array_equal(out[n8:].view(np.uint32), tail[n8:].view(np.uint32))
    assert np.any(lane != tail), "the inputs do not tell the two orders apart"


def _pair_net(rec_type, w=0.5, scale=1.0, n=2):
    """Neuron 0 forced to fire at step 0 (v = 100), synapse 0 -> 1."""
    dt = 0.1
    v0 = np.full(n, -65.0, np.float32)
    v0[0] = 100.0
    pop = lif_pop(n, dt, 0.05, -65.0, -50.0, -65.0, 2.0, v0)
    rp, col, ww, d = csr_from_triplets(n, np.array([0], np.uint32), np.array([1], np.uint32),
                                       np.array([w], np.float32), np.array([0], np.uint8))
    spec = NetSpec("pair", [pop], rp, col, ww, d, dt, 10)
    spec.receptors = dict(types=np.array([rec_type], np.uint8), scale=scale)
    return spec


@pytest.mark.parametrize("code,row", [(0, 0), (1, 1), (4, 4), (6, 6), (9, 0), (200, 0)])
def test_jump_timing_and_closed_form_decay(code, row):
    # FP64 twin: the jump lands after step 0's update (g = w), then decays by
    # exp(-dt/tau) before each later step's current; unknown codes are AMPA
    spec = _pair_net(code, w=0.5)
    ora = po.OracleNetwork(spec, 64)
    ora.step(1)
    assert ora.active().tolist() == [0]
    g = ora.conductances()
    assert g[row, 1] == 0.5 and np.count_nonzero(g) == 1
    ring = ora.ring()
    assert not np.any(ring), "a conductance synapse must not write the ring"
    ora.step(5)
    g = ora.conductances()
    dt = float(np.float32(0.1))  # the population's dt parameter is FP32
    assert abs(g[row, 1] - 0.5 * np.exp(-5 * dt / TAU[row])) < 1e-12


def test_current_based_code_and_scale():
    # 255 keeps the synapse current-based (the ring); the jump scales by `scale`
    ora = po.OracleNetwork(_pair_net(255, w=0.5), 64)
    ora.step(1)
    assert not np.any(ora.conductances())
    assert ora.ring()[1, 1] == 0.5
    ora = po.OracleNetwork(_pair_net(2, w=0.5, scale=3.0), 64)
    ora.step(1)
    assert ora.conductances()[2, 1] == 1.5


@pytest.mark.parametrize("w", [1.0, 2.0])
def test_receptor_current_enters_next_update(w):
    # step 1 of neuron 1: g decays, then I = g f (0 - v) for AMPA at v = -65,
    # clamped to +-100 (population.rs:79-88), and the LIF update (lif.rs:
    # v += ((v_rest - v) + I) inv_tau dt) takes it; FP64 with FP32 parameters
    dt, inv_tau, v_rest = float(np.float32(0.1)), float(np.float32(0.05)), -65.0
    spec = _pair_net(0, w=w)
    ora = po.OracleNetwork(spec, 64)
    ora.step(1)
    v1 = ora.state(0, 0)[1]
    assert v1 == -65.0
    ora.step(1)
    g = w * np.exp(-dt / 2.0)
    i_syn = min(g * (0.0 - v1), 100.0)
    v_exp = v1 + ((v_rest - v1) + i_syn) * inv_tau * dt
    assert abs(ora.state(0, 0)[1] - v_exp) < 1e-9
    assert abs(ora.conductances()[0, 1] - g) < 1e-15


def test_oracle_set_state_resumes_exactly(tmp_path):
    # a state written as a `.taas` snapshot (the layout the device's
    # save_state writes) and loaded into a fresh FP32 twin continues exactly
    # as the twin it came from
    from nets import random_net
    from simuse.formats import TaasWriter
    spec = random_net(600, 20, seed=4, model="lif", max_delay=1, sources=2)
    spec.receptors = dict(types=np.full(spec.n_syn, 0, np.uint8), scale=1.0)
    spec.w = np.abs(spec.w) * np.float32(0.02)
    a = po.OracleNetwork(spec, 32)
    a.step(60)
    w = TaasWriter()
    w.u64("step", [a.step_counter()])
    pop = spec.pops[0]
    w.f32("pop0.states", np.concatenate([a.state(0, c).astype(np.float32) for c in range(pop.ns)]))
    w.f32("ring", a.ring().astype(np.float32).ravel())
    w.f32("weights", a.weights().astype(np.float32))
    w.f32("receptors.g", a.conductances().astype(np.float32).ravel())
    p = tmp_path / "s.taas"
    w.to_file(p)
    b = po.OracleNetwork(spec, 32)
    b.step(7)  # some other state first
    b.set_state_from_snapshot(p)
    assert b.step_counter() == 60
    for _ in range(80):
        a.step()
        b.step()
        assert np.array_equal(a.active(), b.active())
    for c in range(pop.ns):
        assert np.array_equal(a.state(0, c), b.state(0, c))
    assert np.array_equal(a.conductances(), b.conductances())
