"""The alpha-LIF model and the creature pool on the oracle (CPU only).

Pins the oracle's alpha-LIF step and its pool restatement (BrainPool over the
per-post pull scatter, crates/dioma-engine/src/brain/dynamics/gpu_pool.rs and
crates/SiMuse/crates/pillars-wgpu/shaders/csr_scatter.wgsl) with:
* the reference's own alpha-LIF tests (crates/SiMuse/crates/pillars/src/zoo/
  neurons/alpha_lif.rs:139-300) restated, plus a sharper strict-threshold case;
* the fused multiply-adds against exact rational arithmetic, on inputs where
  a separate multiply and add would round differently;
* closed forms for one spike's arrival time, sign and value, for a weight
  change between ticks, and for the once-per-submit voltage jump;
* the pool's pull scatter against the oracle network's push scatter (an
  independent algorithm) on random multi-creature pools, with one delay and
  with per-synapse delays.
"""
from fractions import Fraction

import numpy as np
import pytest

import pyoracle as po
from simuse.models import alpha_lif_pop
from simuse.workloads import NetSpec, creature, pool_net_spec

f32 = np.float32


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _single(n, dt=0.1, tau_m=20.0, tau_ref=2.2, tau_syn=5.0):
    """make_pop of alpha_lif.rs:146-160 (dt 0.1, tf = dt/20, refr 22, syn_decay 1 - dt/5)."""
    pop = alpha_lif_pop(n, dt=dt, tau_ref=tau_ref, tau_m=tau_m, tau_syn=tau_syn)
    rp = np.zeros(n + 1, np.uint32)
    return NetSpec("alif", [pop], rp, np.zeros(0, np.uint32), np.zeros(0, np.float32), np.zeros(0, np.uint8), dt, 0)


def _drive(ora, n, val, steps):
    """Constant input: inject_current puts val into the slot the next step reads."""
    total = 0
    for _ in range(steps):
        ora.inject_current(0, np.arange(n), np.full(n, val, np.float32))
        ora.step()
        total += int(ora.spikes().sum())
    return total


def test_alpha_lif_parameters_match_ensure_built():
    pop = alpha_lif_pop(4)  # the pool's configuration
    assert pop.params[0, 0] == f32(0.5)
    assert pop.params[1, 0] == f32(0.5) / f32(20.0)          # time_factor
    assert pop.params[5, 0] == f32(4.0)                       # max(round(2.2/0.5), 1)
    assert pop.params[6, 0] == f32(1.0) - f32(0.5) / f32(3.0)  # syn_decay
    assert list(pop.states[:, 0]) == [f32(-52.0), f32(0.0), f32(4.0)]
    # f32::round takes halves away from zero: 2.5 -> 3 (numpy would give 2)
    assert alpha_lif_pop(1, dt=1.0, tau_ref=2.5).params[5, 0] == f32(3.0)


def test_alpha_lif_spikes_with_conductance():
    ora = po.OracleNetwork(_single(8), 32)
    assert _drive(ora, 8, 50.0, 200) > 0


def test_alpha_lif_silent_without_input():
    ora = po.OracleNetwork(_single(8), 32)
    assert _drive(ora, 8, 0.0, 500) == 0
    v = ora.state(0, 0)
    assert np.all(v == -52.0)


def test_alpha_lif_strict_threshold_and_old_g():
    # the reference's case: input 1400 at rest cannot spike on its first step
    # (v integrates the OLD g = 0) ...
    ora = po.OracleNetwork(_single(8), 32)
    assert _drive(ora, 8, 1400.0, 1) == 0
    g = ora.state(0, 1)
    assert np.all(g == 1400.0)
    # ... and on the next, g = 1400 carries v from -52 to exactly v_thresh:
    # fma(tf, 1400, -52) = -45 in FP32, which must not spike (strict >)
    assert _drive(ora, 8, 0.0, 1) == 0
    v = ora.state(0, 0)
    assert np.all(v == f32(-45.0))
    # one ulp more of g fires
    ora2 = po.OracleNetwork(_single(1), 32)
    _drive(ora2, 1, float(np.nextafter(f32(1400.0), f32(2000.0))) * 1.0001, 1)
    assert _drive(ora2, 1, 0.0, 1) == 1


def test_alpha_lif_refractory_masks_input():
    ora = po.OracleNetwork(_single(8), 32)
    first = None
    for t in range(50):
        if _drive(ora, 8, 1000.0, 1):
            first = t
            break
    assert first is not None and first >= 1
    assert _drive(ora, 8, 1000.0, 20) == 0
    r = ora.state(0, 2)
    assert np.all(r == 20.0)  # refr_cnt counts every step after the spike


def test_alpha_lif_no_clamp():
    for n in (1, 9):
        ora = po.OracleNetwork(_single(n), 32)
        _drive(ora, n, 500.0, 1)
        assert np.all(ora.state(0, 1) == 500.0)


def test_alpha_lif_tail_parity():
    ora = po.OracleNetwork(_single(13), 32)
    for _ in range(100):
        _drive(ora, 13, 30.0, 1)
        v, g = _bits(ora.state(0, 0)), _bits(ora.state(0, 1))
        assert np.all(v == v[0]) and np.all(g == g[0])


def _round_f32(q: Fraction) -> np.float32:
    """Round an exact rational to the nearest FP32, ties to even."""
    x = np.float32(float(q))
    lo, hi = np.nextafter(x, np.float32(-np.inf)), np.nextafter(x, np.float32(np.inf))
    best = min((abs(Fraction(float(c)) - q), int(np.float32(c).view(np.uint32)) & 1, c) for c in (lo, x, hi))
    return np.float32(best[2])


def test_alpha_lif_fused_multiply_adds_exact():
    """v_new = fma(tf, g - (v - v_rest), v) and g_new = fma(syn_decay, g, I)
    against exact rationals; cases where mul-then-add rounds differently
    make up part of the set, so a dropped fma fails."""
    rng = np.random.default_rng(5)
    n = 4000
    spec = _single(n, dt=0.5, tau_m=20.0, tau_syn=3.0)
    v0 = rng.uniform(-60.0, -45.5, n).astype(np.float32)
    g0 = rng.uniform(-20.0, 40.0, n).astype(np.float32)
    spec.pops[0].states[0, :n] = v0
    spec.pops[0].states[1, :n] = g0
    I = rng.uniform(-30.0, 30.0, n).astype(np.float32)
    ora = po.OracleNetwork(spec, 32)
    ora.inject_current(0, np.arange(n), I)
    ora.step()
    v1, g1 = ora.state(0, 0).astype(np.float32), ora.state(0, 1).astype(np.float32)
    p = spec.pops[0].params[:, 0]
    tf, v_rest, v_th, sd = p[1], p[2], p[3], p[6]
    differs = 0
    for i in range(n):
        integrand = np.float32(g0[i] - np.float32(v0[i] - v_rest))
        vf = _round_f32(Fraction(float(tf)) * Fraction(float(integrand)) + Fraction(float(v0[i])))
        gf = _round_f32(Fraction(float(sd)) * Fraction(float(g0[i])) + Fraction(float(I[i])))
        spike = vf > v_th
        assert _bits(v1[i]) == _bits(f32(-52.0) if spike else vf), i
        assert _bits(g1[i]) == _bits(f32(0.0) if spike else gf), i
        differs += int(np.float32(np.float32(tf * integrand) + v0[i]) != vf)
        differs += int(np.float32(np.float32(sd * g0[i]) + I[i]) != gf)
    assert differs > 100


# ---------------------------------------------------------------------------
# the pool
# ---------------------------------------------------------------------------
def _chain(w=30.0, sign=1.0):
    """0 -> 1 with weight w; neuron 0's transmitter sign."""
    rp = np.array([0, 1, 1], np.int32)
    return rp, np.array([1], np.int32), np.array([w], np.float32), np.array([sign, 1.0], np.float32)


def _g_trace(pool, exts, ticks):
    out = []
    for t in range(ticks):
        pool.step(exts if t == 0 else [None], 1)
        out.append(pool.state(1)[1])
    return np.array(out, np.float32)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_pool_spike_arrival_closed_form(sign):
    """Neuron 0 jumped over threshold fires at step 0; the pull at step 1 puts
    w*sign into slot 1 + d (d = round(1.8/0.5) + 1 = 5), which neuron 1 reads
    at step 6: its g is 0 before and w*sign right after."""
    pool = po.OraclePool()
    assert pool.delay == 5
    pool.register(*_chain(30.0, sign))
    g = _g_trace(pool, [np.array([100.0, 0.0], np.float32)], 10)
    assert np.all(g[:6] == 0.0)
    assert g[6] == f32(30.0 * sign)
    assert g[7] == f32(1.0 - 0.5 / 3.0) * f32(30.0 * sign)


@pytest.mark.parametrize("delay, arrive", [(0, 2), (1, 2), (3, 4), (31, 32)])
def test_pool_delay_override_and_floor(delay, arrive):
    """delay_steps overrides the derived delay; max(d, 1) floors it; 31 is the cap."""
    pool = po.OraclePool(delay_steps=delay)
    pool.register(*_chain())
    g = _g_trace(pool, [np.array([100.0, 0.0], np.float32)], 34)
    assert np.all(g[:arrive] == 0.0) and g[arrive] == f32(30.0)
    assert po.OraclePool(delay_steps=40).delay == 31


def test_pool_weight_sync_reaches_last_substeps_spikes():
    """A weight sync after the tick in which neuron 0 fired still reaches
    that spike (the pull happens at the next step, with the live weights);
    a sync one tick later does not."""
    for sync_at, expect in ((1, 7.0), (2, 30.0)):
        pool = po.OraclePool()
        pool.register(*_chain(30.0))
        g = []
        for t in range(10):
            if t == sync_at:
                pool.set_weights(0, np.array([7.0], np.float32))
            pool.step([np.array([100.0, 0.0], np.float32)] if t == 0 else [None], 1)
            g.append(pool.state(1)[1])
        assert g[6] == f32(expect), sync_at


def test_pool_voltage_jump_once_per_submit():
    """ext is added before the first substep of every submit of <= 64: one
    130-substep tick equals ticks of 64, 64 and 2 with the same input, and
    differs from 130 one-substep ticks."""
    c = creature(60, seed=3)
    ext = np.random.default_rng(1).uniform(0.0, 4.0, 60).astype(np.float32)
    a, b, s = po.OraclePool(), po.OraclePool(), po.OraclePool()
    for p in (a, b, s):
        p.register(c.row_ptr, c.col, c.values, c.nt)
    fa = a.step([ext], 130)
    fb = np.concatenate([b.step([ext], 64), b.step([ext], 64), b.step([ext], 2)])
    fs = np.concatenate([s.step([ext], 1) for _ in range(130)])
    assert np.array_equal(fa, fb)
    assert np.array_equal(_bits(a.state(0)), _bits(b.state(0)))
    assert not np.array_equal(_bits(a.state(0)), _bits(s.state(0)))
    assert fs.sum() > fa.sum()


def test_pool_register_restarts_every_creature():
    c1, c2 = creature(40, seed=1), creature(30, seed=2)
    pool = po.OraclePool()
    pool.register(c1.row_ptr, c1.col, c1.values, c1.nt)
    pool.step([np.full(40, 20.0, np.float32)], 50)
    assert np.any(pool.state(0) != -52.0)
    pool.register(c2.row_ptr, c2.col, c2.values, c2.nt)
    assert np.all(pool.state(0) == -52.0) and np.all(pool.state(1) == 0.0) and np.all(pool.state(2) == 4.0)


def _pool_vs_net(creatures, ticks, spt, exts_fn, delays=None, delay=5):
    pool = po.OraclePool(delay_steps=delay)
    for k, c in enumerate(creatures):
        pool.register(c.row_ptr, c.col, c.values, c.nt)
    if delays is not None:
        for k, d in enumerate(delays):
            pool.set_delays(k, d)
    spec = pool_net_spec(creatures, delay, delays)
    net = po.OracleNetwork(spec, 32)
    n = spec.n_total
    total = 0
    for t in range(ticks):
        exts = exts_fn(t)
        flat = pool.step(exts, spt).reshape(spt, n)
        full = np.zeros(n, np.float32)
        off = 0
        for c, e in zip(creatures, exts):
            if e is not None:
                full[off: off + c.n] = e
            off += c.n
        for k in range(spt):
            if k % 64 == 0:
                net.inject_voltage(0, None, full)
            net.step()
            assert np.array_equal(flat[k].astype(np.int32), net.spikes()), (t, k)
        total += int(flat.sum())
    for col in range(3):
        assert np.array_equal(_bits(pool.state(col)), _bits(net.state(0, col).astype(np.float32))), col
    return total


def _exts(creatures, seed):
    rng = np.random.default_rng(seed)

    def f(t):
        return [rng.uniform(0.0, 9.0, c.n).astype(np.float32) if (t + k) % 3 else None
                for k, c in enumerate(creatures)]
    return f


def test_pool_pull_scatter_equals_push_network():
    cs = [creature(n, p=0.1, seed=s) for s, n in enumerate((50, 120, 7, 200, 33))]
    total = _pool_vs_net(cs, ticks=12, spt=50, exts_fn=_exts(cs, 4))
    assert total > 500  # busy enough that the scatter carries real traffic


def test_pool_pull_scatter_per_synapse_delays_equal_push_network():
    """With several delays a ring cell collects spikes of several steps: the
    pull adds each step's bucket sum, the push adds edge by edge, so the two
    round alike only when the sums are exact.  Integer weights make them
    exact, which leaves the bucketing itself (slot, order, floor) under test."""
    cs = [creature(n, p=0.12, seed=10 + s) for s, n in enumerate((80, 150, 64))]
    for c in cs:
        c.values[:] = np.round(c.values)
    rng = np.random.default_rng(9)
    delays = [rng.integers(0, 31, c.n_syn).astype(np.uint8) for c in cs]
    total = _pool_vs_net(cs, ticks=8, spt=70, exts_fn=_exts(cs, 5), delays=delays)
    assert total > 300


def test_pool_mutant_dropped_sign_is_caught():
    """A planted mistake (transmitter signs ignored) makes the pool differ
    from the push network: the comparison above has teeth for the sign."""
    cs = [creature(120, p=0.1, seed=2)]
    pos = [type(cs[0])(cs[0].row_ptr, cs[0].col, cs[0].values, np.abs(cs[0].nt))]
    pool = po.OraclePool()
    pool.register(pos[0].row_ptr, pos[0].col, pos[0].values, pos[0].nt)
    net = po.OracleNetwork(pool_net_spec(cs, 5), 32)
    ext = np.full(120, 10.0, np.float32)  # every neuron fires at step 0
    flat = pool.step([ext], 60).reshape(60, 120)
    same = True
    for k in range(60):
        if k == 0:
            net.inject_voltage(0, None, ext)
        net.step()
        same &= bool(np.array_equal(flat[k].astype(np.int32), net.spikes()))
    assert not same
