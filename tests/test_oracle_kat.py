"""Pins the CPU oracle against the reference's own known-answer tests and
against closed forms (CPU only).

Reference paths are relative to crates/SiMuse/crates/.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from nets import delay_net, empty_csr
from simuse.models import hh_pop, izh_pop, lif_pop
from simuse.workloads import NetSpec, csr_from_triplets


def single_pop(pop, dt):
    rp, col, w, d = empty_csr(pop.n)
    return NetSpec("single", [pop], rp, col, w, d, dt, 0)


# --- compaction: pillars/src/api/net/compact.rs:79-140 -----------------------
def test_compact_known_answers():
    assert list(po.compact(np.zeros(100, np.int32))) == []
    s = np.array([1 if i % 3 == 0 else 0 for i in range(20)], np.int32)
    assert list(po.compact(s)) == [0, 3, 6, 9, 12, 15, 18]
    assert list(po.compact(np.ones(10, np.int32))) == list(range(10))
    assert list(po.compact(np.array([0, 1] * 5, np.int32))) == [1, 3, 5, 7, 9]
    n = 1003
    i = np.arange(n, dtype=np.uint64)
    s = (((i * np.uint64(2654435761)) % np.uint64(997)) < 97).astype(np.int32)
    assert np.array_equal(po.compact(s), np.nonzero(s)[0].astype(np.uint32))


# --- push / delay ring: pillars/src/api/net/push.rs:300-431 ------------------
def _push_net(n, pre, post, w, d):
    pops = [lif_pop(n, 0.5, 0.5, -70.0, -54.0, -70.0, 2.0, -70.0)]
    rp, col, ww, dd = csr_from_triplets(n, pre, post, np.asarray(w, np.float32), np.asarray(d, np.uint8))
    return NetSpec("push", pops, rp, col, ww, dd, 0.5, 0)


def _force_fire(o, pop, idx):
    # clamp(500) = 100; dv = 100*0.5*0.5 = 25 mV from -70 -> -45 >= -54 fires now
    o.inject_current(pop, idx, np.full(len(idx), 500.0, np.float32))


def test_push_delays_to_slots():
    o = po.OracleNetwork(_push_net(6, [0, 1, 2], [3, 4, 5], [1.0, 2.0, 0.5], [0, 1, 5]))
    _force_fire(o, 0, [0, 1, 2])
    o.step()
    assert list(o.active()) == [0, 1, 2]
    ring = o.ring()
    assert ring[1][3] == 1.0 and ring[1][4] == 2.0 and ring[5][5] == 0.5
    assert ring[0][3] == 0.0 and ring[0][4] == 0.0
    nz = {(s, i) for s, i in zip(*np.nonzero(ring))}
    assert nz == {(1, 3), (1, 4), (5, 5)}


def test_push_wraparound_and_step_counter():
    o = po.OracleNetwork(_push_net(2, [0], [1], [1.0], [31]))
    o.step()  # step 0: nothing fires
    _force_fire(o, 0, [0])
    o.step()  # step 1: fires, (1+31)%32 = 0
    assert o.ring()[0][1] == 1.0
    o2 = po.OracleNetwork(_push_net(2, [0], [1], [1.0], [2]))
    o2.step(5)
    _force_fire(o2, 0, [0])
    o2.step()  # step 5: (5+2)%32 = 7
    r = o2.ring()
    assert r[7][1] == 1.0 and r[0][1] == 0.0


def test_push_accumulates():
    o = po.OracleNetwork(_push_net(3, [0, 1], [2, 2], [3.0, 4.0], [0, 0]))
    _force_fire(o, 0, [0, 1])
    o.step()
    assert abs(o.ring()[1][2] - 7.0) < 1e-6


def test_push_spikes_par_matches_sequential():
    # push.rs:433-458: the threaded push (thread-chunked triples, in-order
    # apply) leaves the ring byte-identical to the sequential one.  Here 300
    # neurons fire at once (above the 64-active fallback), their rows collide
    # on shared posts with mixed delays, and the noise is drawn threaded too.
    rng = np.random.default_rng(5)
    n = 400
    pre = np.repeat(np.arange(n, dtype=np.uint32), 30)
    post = rng.integers(0, 50, size=pre.size).astype(np.uint32)
    w = rng.normal(size=pre.size).astype(np.float32)
    d = rng.integers(0, 32, size=pre.size).astype(np.uint8)
    spec = _push_net(n, pre, post, w, d)
    spec.noise.append((0, np.full(n, 3.0, np.float32), 11))
    runs = []
    for threads in (1, 8):
        o = po.OracleNetwork(spec, 32, threads=threads)
        _force_fire(o, 0, np.arange(300))
        o.step()
        assert o.active().size >= 300
        first = o.ring()
        o.step(40)
        runs.append((first, o.ring(), o.state(0, 0), o.counters()))
    for k in range(3):
        assert np.array_equal(runs[0][k].view(np.uint64), runs[1][k].view(np.uint64))
    assert runs[0][3] == runs[1][3]
    # the cells hold sums of several events whose order changes their bits
    assert np.count_nonzero(runs[0][0]) > 1000


def test_inject_same_step_arrival():
    # build.rs:243-275 — current injected before step s drives step s
    o = po.OracleNetwork(_push_net(2, [0], [1], [1.0], [0]))
    o.inject_current(0, [1], [5.0])
    assert o.ring()[0][1] == 5.0
    o.step()
    assert o.ring()[0][1] == 0.0  # read and cleared


# --- delivery through Network::step: pillars/tests/network_delays.rs --------
@pytest.mark.parametrize("n", [4, 80])
def test_network_delays_exact_arrival(n):
    delays = [0, 2, 5, 31] + [0] * (n - 4)
    o = po.OracleNetwork(delay_net(delays))
    raster = []
    for step in range(70):
        o.step()
        s = o.spikes()
        if step == 0:
            assert s[:n].all()
        raster.append(s[n:].copy())
    raster = np.array(raster)
    for i, d in enumerate(delays):
        arrival = max(d, 1)
        want = np.zeros(70, np.int32)
        want[arrival] = 1
        assert np.array_equal(raster[:, i], want), (i, d)
    assert raster.sum() == n


# --- LIF: pillars/tests/lif_vs_neuron.rs:77-145 + closed form ----------------
@pytest.mark.parametrize("i_syn,ref", [(0.0, 0), (30.0, 51), (50.0, 83)])
def test_lif_spike_counts_vs_reference(i_syn, ref):
    dt, t_ms, n = 0.1, 500.0, 8
    o = po.OracleNetwork(single_pop(lif_pop(n, dt, 0.1, -70.0, -54.0, -70.0, 2.0, -70.0), dt))
    on = int(10.0 / dt)
    counts = np.zeros(n)
    for step in range(int(t_ms / dt)):
        if step >= on:
            o.inject_current(0, np.arange(n), np.full(n, i_syn, np.float32))
        o.step()
        counts += o.spikes()
    assert abs(counts.mean() - ref) <= 2.0


def test_lif_first_spike_matches_closed_form():
    # forward Euler with constant I: v_k = v_inf + (v0 - v_inf)(1-a)^k, a = dt/tau
    dt, tau, I = 0.1, 10.0, 30.0
    v0, vth = -70.0, -54.0
    o = po.OracleNetwork(single_pop(lif_pop(1, dt, 1.0 / tau, -70.0, vth, -70.0, 2.0, v0), dt))
    a = dt / tau
    v_inf = -70.0 + I
    k_star = math.ceil(math.log((vth - v_inf) / (v0 - v_inf)) / math.log(1 - a))
    for step in range(200):
        o.inject_current(0, [0], [I])
        o.step()
        if o.spikes()[0]:
            break
    # step index s fires after s+1 updates
    assert abs((step + 1) - k_star) <= 1


# --- Izhikevich: pillars/tests/izhikevich_vs_neuron.rs:31-214 ---------------
@pytest.mark.parametrize("label,a,b,c,d,i_syn,lo,hi,min_cv", [
    ("RS", 0.02, 0.2, -65.0, 8.0, 10.0, 15, 30, 0.0),
    ("FS", 0.1, 0.2, -65.0, 2.0, 30.0, 180, 330, 0.0),
    ("IB", 0.02, 0.2, -55.0, 4.0, 10.0, 20, 55, 0.5),
    ("CH", 0.02, 0.2, -50.0, 2.0, 15.0, 70, 165, 0.3),
])
def test_izhikevich_patterns(label, a, b, c, d, i_syn, lo, hi, min_cv):
    dt, t_ms, n = 0.5, 1000.0, 8
    o = po.OracleNetwork(single_pop(izh_pop(n, dt, a, b, c, d, 2.0, 30.0, c, 0.0), dt))
    on = int(10.0 / dt)
    per = np.zeros(n)
    # ISIs pooled over the population in neuron order within a step, as the
    # reference test computes them (izhikevich_vs_neuron.rs:129-148)
    isi, last, first = [], 0.0, True
    for step in range(int(t_ms / dt)):
        if step >= on:
            o.inject_current(0, np.arange(n), np.full(n, i_syn, np.float32))
        o.step()
        s = o.spikes()
        per += s
        for j in range(n):
            if s[j]:
                t_now = step * dt
                if not first:
                    isi.append(t_now - last)
                last, first = t_now, False
    assert lo <= per.mean() <= hi, (label, per.mean())
    isi = np.array(isi, np.float32)
    cv = float(np.sqrt(np.mean((isi - isi.mean()) ** 2)) / max(isi.mean(), 1e-6)) if isi.size > 2 else 0.0
    assert cv >= min_cv


def test_izhikevich_fp32_tracks_fp64():
    dt = 0.5
    spec = single_pop(izh_pop(4, dt, 0.02, 0.2, -65.0, 8.0, 2.0, 30.0, -65.0, -13.0), dt)
    o32, o64 = po.OracleNetwork(spec, 32), po.OracleNetwork(spec, 64)
    c32 = c64 = 0
    for _ in range(400):
        for o in (o32, o64):
            o.inject_current(0, np.arange(4), np.full(4, 10.0, np.float32))
            o.step()
        c32 += o32.spikes().sum()
        c64 += o64.spikes().sum()
    assert c32 == c64 and c32 > 0


# --- counter-based inputs ------------------------------------------------------
def test_poisson_roll_rate_and_independence():
    # P/src/api/net/network/sources.rs:5-14: u < p with u uniform on 24 bits
    p, T = 0.2, 20000
    for seed in (1, 7, 12345):
        k = sum(po.poisson_spike(seed, 0, t, p) for t in range(T))
        assert abs(k - p * T) < 4 * math.sqrt(T * p * (1 - p))
    a = [po.poisson_spike(1, 0, t, 0.5) for t in range(512)]
    b = [po.poisson_spike(2, 0, t, 0.5) for t in range(512)]
    assert a != b
    assert not any(po.poisson_spike(3, 0, t, 0.0) for t in range(100))
    assert all(po.poisson_spike(3, 0, t, 1.0) for t in range(100))


M64 = (1 << 64) - 1
PHI64 = 0x9E3779B97F4A7C15

# Published outputs of the SplitMix64 generator (Steele, Lea & Flood 2014;
# Vigna's splitmix64.c: x += phi; z = x; ...finaliser...): the first five
# draws after seeding with 1234567, and the first three after seeding with 0.
SPLITMIX_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                    4593380528125082431, 16408922859458223821]
SPLITMIX_0 = [0xE220A8397B1DCDAF, 7960286522194355700, 487617019471545679]


def _fires_exactly_below(h, seed, stream, step):
    """poisson_spike compares u = (h >> 40) / 2^24 with p: it fires for every p
    above u and for none at or below it (the float grid of u is exact)."""
    u = np.float32((h >> 40) / float(1 << 24))
    above = np.nextafter(u, np.float32(2.0))
    below = np.nextafter(u, np.float32(-1.0)) if u > 0 else None
    assert po.poisson_spike(seed, stream, step, float(above)), (seed, stream, step)
    assert not po.poisson_spike(seed, stream, step, float(u)), (seed, stream, step)
    if below is not None:
        assert not po.poisson_spike(seed, stream, step, float(below)), (seed, stream, step)


def test_poisson_hash_is_splitmix64_on_published_vectors():
    # P/src/api/net/network/sources.rs:5-14: key = (seed << 32) | stream,
    # z = key ^ step*phi, then SplitMix64's step (z += phi, finaliser).  The
    # generator's n-th draw from seed x is finaliser(x + n*phi), so a key
    # k = (x + (n-1)*phi) ^ step*phi hashes to that published draw.  Seed and
    # stream are the key's high and low halves, so a wrong shift, a missing
    # OR, a wrong step multiplier or a dropped z += phi all land elsewhere.
    cases = []
    for n, h in enumerate(SPLITMIX_1234567, start=1):
        for step in (0, 1, 2, 12345, 2 ** 40 + 3):
            z = (1234567 + (n - 1) * PHI64) & M64
            key = z ^ ((step * PHI64) & M64)
            cases.append((h, key >> 32, key & 0xFFFFFFFF, step))
    # seed 0, stream 0: step s hashes s*phi, i.e. the (s+1)-th draw from seed 0
    for s, h in enumerate(SPLITMIX_0):
        cases.append((h, 0, 0, s))
    assert any(c[1] != 0 for c in cases) and any(c[2] != 0 for c in cases)
    for h, seed, stream, step in cases:
        _fires_exactly_below(h, seed, stream, step)


def test_poisson_threshold_on_the_2_pow_minus_24_grid():
    # u takes the top 24 bits of the hash: p exactly on the grid value k/2^24
    # fires only when the draw is below k; p = 1 always fires, p = 0 never
    h = SPLITMIX_0[0]  # seed 0, stream 0, step 0
    k = h >> 40
    assert k == 0xE220A8
    assert not po.poisson_spike(0, 0, 0, k / float(1 << 24))
    assert po.poisson_spike(0, 0, 0, (k + 1) / float(1 << 24))
    assert po.poisson_spike(0, 0, 0, 1.0) and not po.poisson_spike(0, 0, 0, 0.0)


def _splitmix(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_cos2pi_accuracy():
    # every 97th point of the 2^24 grid, plus the quadrant edges
    k = np.concatenate([np.arange(0, 1 << 24, 97), [0, 1 << 22, 1 << 23, 3 << 22, (1 << 24) - 1]])
    u = (k.astype(np.float64) / (1 << 24)).astype(np.float32)
    got = np.array([po.cos2pi_f32(float(x)) for x in u[::7]])
    want = np.cos(2.0 * np.pi * u[::7].astype(np.float64))
    assert np.max(np.abs(got - want)) < 3e-7
    assert po.cos2pi_f32(0.0) == 1.0 and po.cos2pi_f32(0.5) == -1.0
    assert abs(po.cos2pi_f32(0.25)) < 1e-7 and abs(po.cos2pi_f32(0.75)) < 1e-7


def test_gauss_is_box_muller_of_the_counter_hash():
    # z = sqrt(-2 ln u1) cos(2 pi u2) (bit 10 of i clear) or sin(2 pi u2) (set) with
    # u1 = (h>>40 + 1)/2^24, u2 = ((h>>16)&0xFFFFFF)/2^24,
    # h = SplitMix64(SplitMix64(seed ^ p*C1) ^ step*C2), p = i & ~1024: recomputed in float64 here
    for seed, i, t in [(42, 0, 0), (42, 7, 3), (42, 1031, 3), (1, 123456, 999), (1, 123456 ^ 1024, 999),
                       (2 ** 40 + 5, 3, 2 ** 33), (2 ** 40 + 5, 3 + 1024, 2 ** 33)]:
        pair = i & ~1024
        base = _splitmix((seed ^ (pair * 0xD1B54A32D192ED03)) & M64)
        h = _splitmix(base ^ ((t * 0xA24BAED4963EE407) & M64))
        u1 = ((h >> 40) + 1) / 2.0 ** 24
        u2 = ((h >> 16) & 0xFFFFFF) / 2.0 ** 24
        trig = math.sin if i & 1024 else math.cos
        want = math.sqrt(-2.0 * math.log(u1)) * trig(2.0 * math.pi * u2)
        assert abs(po.gauss(seed, i, t) - want) < 2e-6 * max(1.0, abs(want)), (seed, i, t)


def test_sin2pi_accuracy_and_pair_independence():
    k = np.concatenate([np.arange(0, 1 << 24, 97), [0, 1 << 22, 1 << 23, 3 << 22, (1 << 24) - 1]])
    u = (k.astype(np.float64) / (1 << 24)).astype(np.float32)
    got = np.array([po.sin2pi_f32(float(x)) for x in u[::7]])
    want = np.sin(2.0 * np.pi * u[::7].astype(np.float64))
    assert np.max(np.abs(got - want)) < 3e-7
    assert po.sin2pi_f32(0.25) == 1.0 and po.sin2pi_f32(0.75) == -1.0 and po.sin2pi_f32(0.0) == 0.0
    # the two members of a pair are uncorrelated normals
    a = np.array([po.gauss(9, i, t) for i in range(100) for t in range(400)])
    b = np.array([po.gauss(9, i + 1024, t) for i in range(100) for t in range(400)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.02
    assert abs(b.var() - 1.0) < 0.02


def test_gauss_moments():
    z = np.array([po.gauss(42, i, t) for i in list(range(100)) + list(range(1024, 1124)) for t in range(500)])
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.015
    assert abs(np.mean(np.abs(z) > 2.0) - 0.0455) < 0.004
    kurt = np.mean(z ** 4) / z.var() ** 2
    assert abs(kurt - 3.0) < 0.1


def test_exp_ln_accuracy():
    xs = np.linspace(-80, 80, 20001).astype(np.float32)
    got = np.array([po.exp_f32(float(x)) for x in xs])
    want = np.exp(xs.astype(np.float64))
    assert np.max(np.abs(got - want) / want) < 4e-7
    ys = np.geomspace(1e-12, 1.0, 20001).astype(np.float32)
    got = np.array([po.ln_f32(float(y)) for y in ys])
    want = np.log(ys.astype(np.float64))
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1.0)) < 4e-7


def test_vtrap_known_answers_and_window():
    # P/src/math/vtrap.rs:73-117
    assert abs(po.vtrap_f32(0.0, 10.0) - 10.0) < 0.01
    assert abs(po.vtrap_f32(10.0, 1.0) - 10.0 / (math.exp(10.0) - 1.0)) < 1e-3
    assert abs(po.vtrap_f32(-10.0, 1.0) - (-10.0) / (math.exp(-10.0) - 1.0)) < 1e-3
    for x in (-0.01, -0.001, 0.0, 0.001, 0.01):
        assert math.isfinite(po.vtrap_f32(x, 10.0))
    # FP32 value vs FP64 formula across the 0/0 window and outside it
    for x in np.linspace(-60, 60, 4001):
        x = float(np.float32(x))
        want = po.vtrap_f64(x, 10.0)
        assert abs(po.vtrap_f32(x, 10.0) - want) <= 2e-6 * max(abs(want), 1.0)


# --- STDP: pillars/src/ops/plasticity/stdp.rs:182-216 + closed form --------
def _stdp_pair(k, pre_first, tau=20.0, a_plus=0.01, a_minus=0.012, w0=0.5):
    dt = 0.5
    pops = [lif_pop(2, dt, 0.5, -70.0, -54.0, -70.0, 0.0, [100.0 if pre_first else -70.0,
                                                           -70.0 if pre_first else 100.0])]
    rp, col, w, d = csr_from_triplets(2, [0], [1], np.array([w0], np.float32), np.array([5], np.uint8))
    spec = NetSpec("stdp", pops, rp, col, w, d, dt, 0)
    spec.stdp = dict(a_plus=a_plus, a_minus=a_minus, tau_plus=tau, tau_minus=tau, w_min=0.0, w_max=1.0)
    o = po.OracleNetwork(spec)
    o.step()  # step 0: the forced neuron fires
    o.step(k - 1)
    _force_fire(o, 0, [1 if pre_first else 0])
    o.step()  # step k: the partner fires
    return o.weights()[0]


@pytest.mark.parametrize("k", [1, 3, 10])
def test_stdp_pair_closed_form(k):
    dt, tau = 0.5, 20.0
    dec = float(np.float32(math.exp(-dt / tau)))
    assert abs(_stdp_pair(k, True) - (0.5 + 0.01 * dec ** k)) < 1e-6      # pre then post: LTP
    assert abs(_stdp_pair(k, False) - (0.5 - 0.012 * dec ** k)) < 1e-6    # post then pre: LTD


def test_stdp_sign_known_answers():
    # stdp.rs:182-216: lr 0.1, trace 0.8 -> 0.5 becomes 0.58 (LTP) / 0.42 (LTD)
    tau = 0.5 / -math.log(0.8)
    assert abs(_stdp_pair(1, True, tau=tau, a_plus=0.1) - 0.58) < 1e-6
    assert abs(_stdp_pair(1, False, tau=tau, a_minus=0.1) - 0.42) < 1e-6


# --- Tsodyks-Markram short-term plasticity: pillars/src/api/net/push.rs:515-695 ---------
REF_DEP = [0.500000, 0.295317, 0.211527, 0.177227, 0.163185, 0.157437, 0.155084, 0.154121,
           0.153726, 0.153565, 0.153499, 0.153472]
REF_FAC = [0.100000, 0.144296, 0.148667, 0.132751, 0.110006, 0.087407, 0.067775, 0.051800,
           0.039252, 0.029592, 0.022241, 0.016687]


def _f32exp(x):
    return float(np.exp(np.float32(x)))


def test_tm_matches_neuron_delivered_weights():
    # NEURON tm.mod golden vectors (push.rs:646-695): spikes at steps 20 + 40k, dt 0.5 ms
    dep = [1.0, 0.5, 0]
    fac = [1.0, 0.1, 0]
    for k in range(12):
        step = 20 + 40 * k
        w = po.tm_spike(0.5, _f32exp(-0.5 / 100.0), 0.0, False, dep, 1.0, step)
        assert abs(w - REF_DEP[k]) < 1e-4, (k, w)
        w = po.tm_spike(0.1, _f32exp(-0.5 / 1e6), _f32exp(-0.5 / 50.0), True, fac, 1.0, step)
        assert abs(w - REF_FAC[k]) < 1e-4, (k, w)


def test_tm_depression_recovery_and_facilitation_rules():
    # push.rs:519-543: negligible recovery -> 0.5, 0.25, 0.125, 0.0625
    st = [1.0, 0.5, 0]
    a = _f32exp(-0.5e-6 / 1e6)
    got = [po.tm_spike(0.5, a, 0.0, False, st, 1.0, s) for s in (0, 10, 20, 30)]
    assert np.allclose(got, [0.5, 0.25, 0.125, 0.0625], atol=1e-3)
    # push.rs:546-588: spaced spikes recover to ~U0, tight ones keep decaying
    t, s = [1.0, 0.5, 0], [1.0, 0.5, 0]
    tight = [po.tm_spike(0.5, _f32exp(-1.0), 0.0, False, t, 1.0, k) for k in range(4)]
    spaced = [po.tm_spike(0.5, _f32exp(-1.0), 0.0, False, s, 1.0, k) for k in (0, 1000, 2000, 3000)]
    assert abs(tight[0] - spaced[0]) < 1e-4
    assert all(spaced[i] > tight[i] and abs(spaced[i] - 0.5) < 0.05 for i in range(1, 4))
    # push.rs:591-612: facilitation, U 0.1 -> 0.19 -> ~0.2198
    f = [1.0, 0.1, 0]
    w0 = po.tm_spike(0.1, _f32exp(-1e-6), _f32exp(-1.0), True, f, 1.0, 0)
    w1 = po.tm_spike(0.1, _f32exp(-1e-6), _f32exp(-1.0), True, f, 1.0, 1)
    assert w1 > w0 and abs(f[1] - 0.2198) < 1e-3
    # gap capped at 1e6 steps, recovery exact at a_rec = 1 (closed form r = (1-U)^k)
    g = [1.0, 0.5, 0]
    ws = [po.tm_spike(0.5, 1.0, 0.0, False, g, 2.0, k) for k in range(5)]
    assert ws == [2.0 * 0.5 * 0.5 ** k for k in range(5)]


def test_powi_is_binary_exponentiation():
    # exact for powers of two; elsewhere within the repeated-squaring bound 2n*2^-24
    assert po.powi_f32(0.5, 64) == 2.0 ** -64 and po.powi_f32(2.0, 100) == 2.0 ** 100
    for a in (0.99, 1.0001, 0.36787945, 0.9999):
        for n in (1, 2, 3, 7, 64, 1000, 12345):
            want = float(np.float64(np.float32(a)) ** n)
            if not (1e-37 < want < 1e37):
                continue
            assert abs(po.powi_f32(a, n) - want) <= 2 * n * 2.0 ** -24 * want, (a, n)


def test_tm_network_depresses_repeated_drive():
    # pillars/src/api/net/network/tests.rs:530-570: pre driven every step; the static edge keeps
    # post 1 firing, the depressing edge (U0 0.5, tau_rec 10 ms) silences post 2
    pop = lif_pop(3, 0.5, 0.1, -70.0, -65.0, -70.0, 1.0, -70.0)
    rp, col, w, d = csr_from_triplets(3, [0, 0], [1, 2], np.array([100.0, 100.0], np.float32),
                                      np.zeros(2, np.uint8))
    spec = NetSpec("tm", [pop], rp, col, w, d, 0.5, 0)
    spec.stp = dict(profiles=[(0.5, 10.0, 0.0)], dt=0.5, syn_types=np.array([0, 1], np.uint8))
    o = po.OracleNetwork(spec)
    fired = np.zeros(3, np.int64)
    for _ in range(500):
        o.inject_current(0, [0], [100.0])
        o.step()
        fired += o.spikes()
    assert fired[1] > 20 and fired[2] == 0


# --- GPEA: pillars/src/zoo/neurons/gpea.rs:430-611 (NEURON 9.0.2 values) -----------
def test_fast_exp_accuracy():
    xs = np.linspace(-87.0, 87.0, 40001).astype(np.float32)
    got = np.array([po.fast_exp(float(x)) for x in xs])
    want = np.exp(xs.astype(np.float64))
    assert np.max(np.abs(got - want) / want) < 2e-5
    assert po.fast_exp(0.0) == 1.0 and po.fast_exp(-200.0) == 0.0 and math.isinf(po.fast_exp(200.0))


def test_gpea_rates_vs_neuron():
    r = po.gpea_rates(-65.0, 1e-4)
    want = [0.02069111227, 0.9546474194, 0.1620766655, 0.2069705933, 0.03082031845, 0.2942149722]
    assert np.all(np.abs(r[:6] - want) < 1e-4)
    assert np.all(np.abs(r[6:10] - [0.1009999964, 1.540321035, 1.098909804, 8.077265928]) < 1e-3)
    assert abs(r[10] - 136.9174254) < 5e-2
    r = po.gpea_rates(0.0, 1e-4)
    assert abs(r[0] - 0.9956298205) < 1e-4 and abs(r[1] - 0.0008167381047) < 1e-4
    assert abs(r[3] - 0.9997655911) < 1e-4 and abs(r[10] - 14.24726609) < 1e-2
    r = po.gpea_rates(-80.0, 5e-4)
    assert abs(r[0] - 0.002472623157) < 1e-4 and abs(r[1] - 0.9954615683) < 1e-4
    assert abs(r[4] - 0.2969072608) < 1e-4 and abs(r[5] - 0.9840936083) < 1e-4
    assert np.all(np.abs(r[7:10] - [0.6760660476, 0.925418813, 6.400175291]) < 1e-3)
    assert abs(r[10] - 60.09475979) < 5e-2


def _gpea_run(steps, current, n=8, kind=32, first_spike_after=None):
    from simuse.models import gpea_pop
    pop = gpea_pop(n)
    rp, col, w, d = csr_from_triplets(n, [], [], np.zeros(0, np.float32), np.zeros(0, np.uint8))
    o = po.OracleNetwork(NetSpec("gpea", [pop], rp, col, w, d, 0.025, 0), kind)
    idx = np.arange(n)
    total, first = 0, None
    for k in range(steps):
        if current:
            o.inject_current(0, idx, np.full(n, current, np.float32))
        o.step()
        s = int(o.spikes().sum())
        total += s
        if s and first is None:
            first = k
    return o, total, first


def test_gpea_resting_and_silence_match_neuron():
    o, total, _ = _gpea_run(200, 0.0)
    assert abs(o.state(0, 0)[0] - -61.72048753) < 0.05          # t = 5 ms
    o, total, _ = _gpea_run(2000, 0.0)
    assert abs(o.state(0, 0)[0] - -61.2831801) < 0.05           # t = 50 ms
    o, total, _ = _gpea_run(8000, 0.0)
    assert abs(o.state(0, 0)[0] - -61.04569455) < 0.05          # t = 200 ms
    assert total == 0 and np.all(np.isfinite(o.state(0, 0)))


def test_gpea_pulse_matches_neuron():
    from simuse.models import GPEA_PULSE, gpea_pop
    n = 8
    pop = gpea_pop(n)
    rp, col, w, d = csr_from_triplets(n, [], [], np.zeros(0, np.float32), np.zeros(0, np.uint8))
    o = po.OracleNetwork(NetSpec("gpea", [pop], rp, col, w, d, 0.025, 0))
    o.step(2000)                                                # 0..50 ms at rest
    idx = np.arange(n)
    first, total = None, 0
    for k in range(4000):                                       # 50..150 ms, 0.5 nA
        o.inject_current(0, idx, np.full(n, GPEA_PULSE, np.float32))
        o.step()
        s = int(o.spikes().sum())
        total += s
        if s and first is None:
            first = 50.0 + (k + 0.5) * 0.025
    assert first is not None and abs(first - 50.8) < 0.5       # NEURON: 50.8 ms
    assert abs(total / n - 37.0) <= 3.0                         # NEURON: 37 spikes


def test_gpea_stress_strong_input():
    o, total, _ = _gpea_run(8000, 500.0)                        # 0.5 mA/cm^2, 200 ms
    v = o.state(0, 0)
    assert np.all(np.isfinite(v)) and np.all(np.abs(v) < 200.0)
    assert abs(total / 8 - 220.0) <= 22.0                       # NEURON: 220 spikes


def test_gpea_fp32_tracks_fp64_truth():
    # the FP64 twin uses exact exp/ln; spike counts under the pulse agree
    from simuse.models import GPEA_PULSE
    _, t32, f32 = _gpea_run(4000, GPEA_PULSE)
    _, t64, f64 = _gpea_run(4000, GPEA_PULSE, kind=64)
    assert abs(t32 - t64) <= 8 and abs(f32 - f64) <= 2


# --- dense STDP of CompiledNetwork::stdp_step (pillars-cuda/src/network.rs:53-83) -------
def _dense_run(k, pre_first, decay=0.9, lr_ltp=0.01, lr_ltd=0.02, w0=0.5, w_max=1.0):
    dt = 0.5
    pops = [lif_pop(2, dt, 0.5, -70.0, -54.0, -70.0, 0.0, [100.0 if pre_first else -70.0,
                                                           -70.0 if pre_first else 100.0])]
    rp, col, w, d = csr_from_triplets(2, [0], [1], np.array([w0], np.float32), np.array([5], np.uint8))
    o = po.OracleNetwork(NetSpec("dense", pops, rp, col, w, d, dt, 0))
    for _ in range(k + 1):  # call 1 is the step in which the forced neuron fires
        o.stdp_step(decay, lr_ltp, lr_ltd, 0.0, w_max)
    return o.weights()[0]


@pytest.mark.parametrize("k", [1, 4, 12])
def test_dense_stdp_closed_form(k):
    # the trace of the neuron that fired at step 0 is 1, dec, dec^2, ... from the next
    # call on (traces see the previous step's spikes); every call adds lr * trace
    dec = 0.9
    geo = (1 - dec ** k) / (1 - dec)
    assert abs(_dense_run(k, True) - (0.5 + 0.01 * geo)) < 1e-6    # pre fired: LTP only
    assert abs(_dense_run(k, False) - (0.5 - 0.02 * geo)) < 1e-6   # post fired: LTD only
    # clipping at w_max
    assert _dense_run(k, True, lr_ltp=1.0) == 1.0


# --- statistics: pillars/src/kernels/stats/{rates,isi}.rs -----------------------
def test_rates_and_cv_brute_force():
    rng = np.random.default_rng(3)
    T, n, dt_s = 3000, 40, 0.001
    r = (rng.random((T, n)) < rng.uniform(0.001, 0.05, n)).astype(np.int32)
    rates = po.firing_rates(r, dt_s)
    assert np.allclose(rates, r.sum(0) / (T * dt_s), rtol=1e-5)
    cvs = po.isi_cv(r, dt_s)
    for i in range(n):
        t = np.nonzero(r[:, i])[0]
        isi = np.diff(t).astype(np.float64)
        want = isi.std(ddof=1) / isi.mean() if isi.size >= 2 else 0.0
        assert abs(cvs[i] - want) < 1e-4 * max(1.0, want)


# --- Kuramoto: tests/kuramoto_tests.rs:22-94 + closed forms -----------------
def _all_to_all(n, k):
    rp = np.arange(n + 1, dtype=np.uint32) * (n - 1)
    col = np.array([j for i in range(n) for j in range(n) if j != i], np.uint32)
    return rp, col, np.full(col.size, k / n, np.float32)


def test_kuramoto_sync_from_zero():
    n = 8
    rp, col, w = _all_to_all(n, 1.0)
    k = po.OracleKuramoto(rp, col, w, np.ones(n, np.float32), np.zeros(n, np.float32), 0.5)
    k.euler(100)
    p = k.phases()
    assert np.all(np.abs(p - p[0]) < 0.5)


def test_kuramoto_identical_frequencies_synchronise():
    n = 64
    rng = np.random.default_rng(1)
    rp, col, w = _all_to_all(n, 2.0)
    k = po.OracleKuramoto(rp, col, w, np.full(n, 0.3, np.float32),
                          rng.uniform(-1.5, 1.5, n).astype(np.float32), 0.05)
    r0 = k.order()
    k.rk4(2000)
    assert k.order() > 0.999 > r0


def test_kuramoto_uncoupled_is_linear():
    n = 16
    rng = np.random.default_rng(2)
    om = rng.normal(size=n).astype(np.float32)
    th = rng.uniform(-3, 3, n).astype(np.float32)
    rp = np.zeros(n + 1, np.uint32)
    k = po.OracleKuramoto(rp, np.zeros(0, np.uint32), np.zeros(0, np.float32), om, th, 0.01)
    k.rk4(1000)
    want = th.astype(np.float64) + om.astype(np.float64) * (1000 * float(np.float32(0.01)))
    diff = np.angle(np.exp(1j * (k.phases() - want)))
    assert np.max(np.abs(diff)) < 1e-9


def test_kuramoto_asymmetric_pair_locks_at_closed_form():
    # Two oscillators, W_01 != W_10 (P/src/kernels/phase/sparse.rs:63-96:
    # c_i = sum over row i of W_ij sin(theta_j - theta_i)).  With phi =
    # theta_1 - theta_0: dphi/dt = (w1 - w0) - (W_01 + W_10) sin(phi), so the
    # pair locks at sin(phi*) = (w1 - w0) / (W_01 + W_10) and turns together
    # at Omega = (W_10 w0 + W_01 w1) / (W_01 + W_10).  Ignoring the weights,
    # swapping them (row vs column), an extra 1/N or a sign error each move
    # phi* or Omega.
    w01, w10 = 0.9, 0.3
    om = np.array([0.2, 0.7], np.float32)
    rp = np.array([0, 1, 2], np.uint32)
    col = np.array([1, 0], np.uint32)
    w = np.array([w01, w10], np.float32)
    dt = 0.01
    k = po.OracleKuramoto(rp, col, w, om, np.array([0.0, 0.0], np.float32), dt)
    k.rk4(4000)  # 40 time units: relaxation rate sqrt(1.2^2 - 0.5^2) ~ 1.09
    p0 = k.phases()
    k.rk4(1000)
    p1 = k.phases()
    w01f, w10f = float(np.float32(w01)), float(np.float32(w10))
    om0, om1 = float(om[0]), float(om[1])
    phi = math.atan2(math.sin(p1[1] - p1[0]), math.cos(p1[1] - p1[0]))
    assert abs(math.sin(phi) - (om1 - om0) / (w01f + w10f)) < 1e-9
    assert abs(math.cos(phi) - math.sqrt(1 - ((om1 - om0) / (w01f + w10f)) ** 2)) < 1e-9  # the stable root
    omega = (w10f * om0 + w01f * om1) / (w01f + w10f)
    adv = np.angle(np.exp(1j * (p1 - p0)))
    want = (omega * 1000 * float(np.float32(dt)) + math.pi) % (2 * math.pi) - math.pi
    assert np.all(np.abs(np.angle(np.exp(1j * (adv - want)))) < 1e-9)


def test_kuramoto_euler_step_brute_force_weighted_graph():
    # One forward-Euler step on a 4-node graph with distinct, signed weights
    # and a multi-edge, from the definition theta_i + dt (w_i + sum_e W_e
    # sin(theta_col[e] - theta_i)), wrapped to [-pi, pi) (kuramoto.rs:42-68,
    # P/src/kernels/mod.rs:74-77).
    rp = np.array([0, 2, 3, 6, 7], np.uint32)
    col = np.array([1, 3, 2, 0, 1, 1, 0], np.uint32)
    w = np.array([0.5, -1.25, 2.0, 0.75, 0.1, 0.3, -0.6], np.float32)
    om = np.array([1.0, -0.5, 0.25, 3.0], np.float32)
    th = np.array([0.3, -2.9, 1.7, 3.1], np.float32)
    dt = 0.25
    k = po.OracleKuramoto(rp, col, w, om, th, dt)
    k.euler(1)
    want = []
    for i in range(4):
        c = sum(float(w[e]) * math.sin(float(th[col[e]]) - float(th[i])) for e in range(rp[i], rp[i + 1]))
        x = float(th[i]) + float(np.float32(dt)) * (float(om[i]) + c)
        want.append(x - 2 * math.pi * round(x / (2 * math.pi)))
    assert np.max(np.abs(k.phases() - np.array(want))) < 1e-12
    # node 3's step crosses +pi and wraps
    assert k.phases()[3] < 0 < float(th[3])


def test_kuramoto_rk4_fourth_order():
    n = 12
    rng = np.random.default_rng(4)
    rp, col, w = _all_to_all(n, 1.5)
    om = rng.normal(size=n).astype(np.float32)
    th = rng.uniform(-3, 3, n).astype(np.float32)

    def run(dt, steps):
        k = po.OracleKuramoto(rp, col, w, om, th, dt)
        k.rk4(steps)
        return k.phases()

    ref = run(0.0025, 800)
    e1 = np.max(np.abs(np.angle(np.exp(1j * (run(0.04, 50) - ref)))))
    e2 = np.max(np.abs(np.angle(np.exp(1j * (run(0.02, 100) - ref)))))
    assert 12.0 < e1 / e2 < 20.0


# --- Hodgkin–Huxley: textbook closed forms -----------------------------------
def test_hh_resting_gates_are_steady_state():
    # m_inf, h_inf, n_inf at -65 mV equal the initial gates the reference uses
    # (pillars-connectome/src/sim/hh.rs:34-38)
    v = -65.0
    am = 0.1 * (v + 40) / (1 - math.exp(-(v + 40) / 10))
    bm = 4 * math.exp(-(v + 65) / 18)
    ah = 0.07 * math.exp(-(v + 65) / 20)
    bh = 1 / (1 + math.exp(-(v + 35) / 10))
    an = 0.01 * (v + 55) / (1 - math.exp(-(v + 55) / 10))
    bn = 0.125 * math.exp(-(v + 65) / 80)
    assert abs(am / (am + bm) - 0.0529) < 1e-3
    assert abs(ah / (ah + bh) - 0.596) < 1e-3
    assert abs(an / (an + bn) - 0.318) < 1e-3
    # with the squid-axon leak reversal the membrane stays at rest
    dt = 0.01
    spec = single_pop(hh_pop(2, dt, e_leak=-54.387, tau_ref=2.0), dt)
    for kind in (32, 64):
        o = po.OracleNetwork(spec, kind)
        o.step(2000)
        assert np.all(np.abs(o.state(0, 0) + 65.0) < 0.5)
        assert o.counters()[0] == 0


def test_hh_repetitive_firing_frequency():
    # textbook HH at 6.3 C fires ~68 Hz for a 10 uA/cm^2 step
    dt = 0.01
    spec = single_pop(hh_pop(1, dt, e_leak=-54.387, tau_ref=1.0), dt)
    res = {}
    for kind in (32, 64):
        o = po.OracleNetwork(spec, kind)
        times = []
        for step in range(50000):
            o.inject_current(0, [0], [10.0])
            o.step()
            if o.spikes()[0]:
                times.append(step)
        res[kind] = times
        isi = np.diff(times[2:]) * dt
        assert 60.0 < 1000.0 / isi.mean() < 76.0
    assert len(res[32]) == len(res[64])
    assert np.max(np.abs(np.array(res[32]) - np.array(res[64]))) <= 2
