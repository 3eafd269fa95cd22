"""Parity of the HIP network step with the CPU oracle (needs an MI355X).

Integer/index results (spike ids, counts, ring slots) and the FP32 state are
compared bit for bit: both sides evaluate the reference's operation sequence
with IEEE float ops and no FMA contraction, and the device scatter applies
contributions in the CPU push order.  Reference paths: crates/SiMuse/crates/.
"""
import sys

import numpy as np
import pytest

import pyoracle as po
from nets import delay_net, empty_csr, random_net, two_pop_net
from simuse import SCATTER_TILE, SCATTER_WAVE
from simuse.models import hh_pop, izh_pop, lif_pop
from simuse.workloads import NetSpec, build_device, csr_from_triplets, izhikevich_network

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_same_state(dev, ora, spec):
    for p, pop in enumerate(spec.pops):
        st = dev.download_states(p)
        for c in range(pop.ns):
            a = st[c, : pop.n]
            b = ora.state(p, c).astype(np.float32)
            bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
            assert bad.size == 0, f"pop {p} col {c}: {bad.size} differ, first {bad[:5]} dev={a[bad[:5]]} ora={b[bad[:5]]}"


def assert_same_ring(dev, ora):
    a = dev.download_ring()
    b = ora.ring().astype(np.float32)
    bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
    assert bad.shape[0] == 0, f"ring differs at {bad[:5].tolist()}"


def run_pair(spec, steps, sched=SCATTER_TILE, inject=None, check_every=1, graph_steps=8, async_steps=False,
             tiles=2):
    dev = build_device(spec, graph_steps=graph_steps)
    dev.set_scatter_mode(sched)
    dev.set_block_tiles(tiles)
    ora = po.OracleNetwork(spec, 32)
    if async_steps and inject is None:
        dev.record(steps)
        ora.record(True)
        dev.step_graph(steps)
        dev.sync()
        ora.step(steps)
        assert np.array_equal(dev.download_raster(), ora.raster())
        return dev, ora
    for t in range(steps):
        if inject is not None:
            inject(dev, ora, t)
        dev.step()
        ora.step()
        if check_every and t % check_every == 0:
            a = dev.download_spike_ids()
            b = ora.active()
            assert np.array_equal(a, b), f"step {t}: spikes differ ({a.size} vs {b.size})"
    return dev, ora


def single_pop(pop, dt):
    rp, col, w, d = empty_csr(pop.n)
    return NetSpec("single", [pop], rp, col, w, d, dt, 0)


def test_lif_population_bitwise():
    # pillars-cuda/tests/lif_parity.rs:26-75: N=512, dt 0.1, I=30 on even / 0 on odd, 500 steps
    say("lif_population_bitwise")
    n = 512
    spec = single_pop(lif_pop(n, 0.1, 0.1, -70.0, -54.0, -70.0, 2.0, -70.0), 0.1)
    idx = np.arange(n)
    cur = np.where(idx % 2 == 0, 30.0, 0.0).astype(np.float32)

    def inj(dev, ora, t):
        dev.inject_current(0, idx, cur)
        ora.inject_current(0, idx, cur)

    dev, ora = run_pair(spec, 500, inject=inj)
    assert ora.counters()[0] > 0
    assert_same_state(dev, ora, spec)


def test_izhikevich_population_bitwise():
    # pillars-cuda/tests/izhikevich_parity.rs:28-104: N=64, dt 0.5, I=10, 500 steps
    say("izhikevich_population_bitwise")
    n = 64
    spec = single_pop(izh_pop(n, 0.5, 0.02, 0.2, -65.0, 8.0, 2.0, 30.0, -65.0, 0.0), 0.5)
    idx = np.arange(n)
    cur = np.full(n, 10.0, np.float32)

    def inj(dev, ora, t):
        dev.inject_current(0, idx, cur)
        ora.inject_current(0, idx, cur)

    dev, ora = run_pair(spec, 500, inject=inj)
    assert ora.counters()[0] > 0
    assert_same_state(dev, ora, spec)


def test_two_population_network():
    # pillars-cuda/tests/network_parity.rs:72-181: LIF(8) <-> Izh(8), drive 30, 200 steps
    say("two_population_network")
    spec = two_pop_net(8)
    idx = np.arange(8)
    cur = np.full(8, 30.0, np.float32)
    counts_d = np.zeros(16, np.int64)
    counts_o = np.zeros(16, np.int64)

    def inj(dev, ora, t):
        dev.inject_current(0, idx, cur)
        ora.inject_current(0, idx, cur)

    dev = build_device(spec)
    ora = po.OracleNetwork(spec, 32)
    for t in range(200):
        inj(dev, ora, t)
        dev.step()
        ora.step()
        counts_d += dev.download_spikes()
        counts_o += ora.spikes()
    assert counts_o.sum() > 0
    assert np.array_equal(counts_d, counts_o)
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


@pytest.mark.parametrize("n", [4, 80])
def test_network_delays_exact_arrival(n):
    # pillars/tests/network_delays.rs:139-158
    say("network_delays", n)
    delays = [0, 2, 5, 31] + [0] * (n - 4)
    spec = delay_net(delays)
    dev = build_device(spec)
    raster = []
    for step in range(70):
        dev.step()
        s = dev.download_spikes()
        if step == 0:
            assert s[:n].all()
        raster.append(s[n:])
    raster = np.array(raster)
    for i, d in enumerate(delays):
        want = np.zeros(70, np.int32)
        want[max(d, 1)] = 1
        assert np.array_equal(raster[:, i], want), (i, d)


@pytest.mark.parametrize("sched,tiles", [(SCATTER_TILE, 2), (SCATTER_WAVE, 2), (SCATTER_TILE, 1),
                                         (SCATTER_WAVE, 1)])
def test_random_izhikevich_network_bitexact(sched, tiles):
    # 10k neurons, ~100 synapses each, multi-edges, noise; 1000 steps
    say("random_izhikevich", sched, tiles)
    spec = random_net(10_000, 100, seed=5, model="izh", max_delay=0)
    dev, ora = run_pair(spec, 1000, sched=sched, check_every=1, tiles=tiles)
    s, e = ora.counters()
    assert s > 1000 and e > 0
    assert dev.counters() == (s, e)
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


@pytest.mark.parametrize("sched", [SCATTER_TILE, SCATTER_WAVE])
def test_lif_delays_sources_bitexact(sched):
    say("lif_delays_sources", sched)
    spec = random_net(3000, 40, seed=11, model="lif", max_delay=31, sources=40)
    dev, ora = run_pair(spec, 600, sched=sched, check_every=1)
    assert ora.counters()[0] > 100
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


@pytest.mark.parametrize("sched", [SCATTER_TILE, SCATTER_WAVE])
def test_unit_delay_network_with_sources_and_injection(sched):
    # every synapse and source delay <= 1: the scatter skips the ring read
    say("unit_delay", sched)
    spec = random_net(3000, 60, seed=13, model="lif", max_delay=1, sources=30)
    idx = np.arange(0, 3000, 7)
    cur = np.linspace(-20.0, 40.0, idx.size).astype(np.float32)

    def inj(dv, o, t):
        if t % 3 == 0:
            dv.inject_current(0, idx, cur)
            o.inject_current(0, idx, cur)

    dev, ora = run_pair(spec, 400, sched=sched, inject=inj)
    assert ora.counters()[0] > 100
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


def test_binned_scatter_oversized_segments_and_hubs():
    # delays <= 1 (binned scatter): every neuron fires every step and sends 40 edges into the
    # first 2048 posts, so one tile puts 10240 records into one bin (more than one LDS chunk)
    # and a few hub posts collect hundreds of events per step
    say("binned_oversized")
    rng = np.random.default_rng(17)
    n, deg = 4096, 40
    pre = np.repeat(np.arange(n, dtype=np.uint32), deg)
    post = np.where(rng.random(pre.size) < 0.1, rng.integers(0, 4, pre.size),
                    rng.integers(0, 2048, pre.size)).astype(np.uint32)
    w = rng.normal(0.0, 1.0, pre.size).astype(np.float32)
    d = rng.integers(0, 2, pre.size).astype(np.uint8)
    rp, col, w, d = csr_from_triplets(n, pre, post, w, d)
    spec = NetSpec("hubs", [lif_pop(n, 0.5, 0.5, -70.0, -54.0, -70.0, 0.0, -70.0)], rp, col, w, d, 0.5, 0)
    idx = np.arange(n)
    big = np.full(n, 500.0, np.float32)
    dev, ora = run_pair(spec, 12, inject=lambda dv, o, t: (dv.inject_current(0, idx, big),
                                                         o.inject_current(0, idx, big)))
    assert ora.counters()[1] >= 12 * n * deg
    assert dev.counters() == ora.counters()
    assert_same_ring(dev, ora)
    assert_same_state(dev, ora, spec)


def test_binned_scatter_many_tiles_with_sources():
    # 80k neurons = 313 neuron tiles + source tiles: several tiles per bin block
    say("binned_many_tiles")
    spec = random_net(80_000, 30, seed=19, model="lif", max_delay=1, sources=300)
    dev, ora = run_pair(spec, 150, check_every=10)
    assert ora.counters()[0] > 1000
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


def test_binned_scatter_multi_chunk_blocks():
    # 1.2M neurons = 4688 neuron tiles, so a bin block covers 5 tiles; with every
    # neuron firing a block stages 1280 spiking rows, more than one LDS chunk
    say("binned_multi_chunk")
    rng = np.random.default_rng(23)
    n, deg = 1_200_000, 3
    pre = np.repeat(np.arange(n, dtype=np.uint32), deg)
    post = rng.integers(0, n, pre.size).astype(np.uint32)
    w = rng.normal(0.0, 1.0, pre.size).astype(np.float32)
    d = np.ones(pre.size, np.uint8)
    rp, col, w, d = csr_from_triplets(n, pre, post, w, d)
    spec = NetSpec("wide", [lif_pop(n, 0.5, 0.5, -70.0, -54.0, -70.0, 0.0, -70.0)], rp, col, w, d, 0.5, 0)
    idx = np.arange(n, dtype=np.uint32)
    big = np.full(n, 500.0, np.float32)
    dev, ora = run_pair(spec, 3, inject=lambda dv, o, t: (dv.inject_current(0, idx, big),
                                                        o.inject_current(0, idx, big)))
    assert ora.counters()[0] >= 2 * n
    assert dev.counters() == ora.counters()
    assert_same_ring(dev, ora)
    assert_same_state(dev, ora, spec)


def test_graph_replay_matches_stepwise():
    say("graph_replay")
    spec = random_net(5000, 60, seed=3, model="izh", max_delay=7, sources=5)
    for gs in (1, 5, 16):
        dev, ora = run_pair(spec, 203, graph_steps=gs, async_steps=True)
        assert_same_state(dev, ora, spec)
        assert_same_ring(dev, ora)


def test_brunel_lif_small_bitexact():
    # the Brunel LIF benchmark network at reduced size: delays 1..15, Poisson drive per neuron
    from simuse.workloads import brunel_lif_network
    say("brunel_small")
    spec = brunel_lif_network(n=2000, p=0.05, seed=2, steps=1500, rate_hz=20000.0)
    dev, ora = run_pair(spec, 1500, async_steps=True)
    assert ora.counters()[0] > 0
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


def test_hh_network_bitexact_and_fp64_error():
    say("hh_network")
    spec = random_net(2000, 50, seed=7, model="hh", max_delay=3, noise=40.0)
    dev = build_device(spec)
    o32 = po.OracleNetwork(spec, 32)
    o64 = po.OracleNetwork(spec, 64)
    dev.record(1500)
    o32.record(True)
    dev.run(300)
    o32.step(300)
    o64.step(300)
    v_dev = dev.download_states(0)[0, :2000]
    v64 = o64.state(0, 0)
    # short-window error of FP32 state against the FP64 truth
    rel = np.max(np.abs(v_dev - v64)) / np.max(np.abs(v64))
    say("hh max rel v error vs fp64 after 300 steps", rel)
    assert rel < 1e-3
    dev.run(1200)
    o32.step(1200)
    assert np.array_equal(dev.download_raster(), o32.raster())
    assert o32.counters()[0] > 0
    assert_same_state(dev, o32, spec)
    assert_same_ring(dev, o32)


def test_stdp_weights_bitexact():
    say("stdp")
    spec = random_net(3000, 50, seed=13, model="lif", max_delay=4, sources=10)
    rng = np.random.default_rng(0)
    plastic = (rng.random(spec.n_syn) < 0.7).astype(np.uint8)
    spec.w = np.abs(spec.w) / 40.0
    spec.stdp = dict(a_plus=0.01, a_minus=0.012, tau_plus=20.0, tau_minus=20.0, w_min=0.0, w_max=1.0,
                     plastic=plastic)
    dev, ora = run_pair(spec, 400, check_every=1)
    assert ora.counters()[0] > 100
    wd = dev.download_weights()
    wo = ora.weights().astype(np.float32)
    assert np.array_equal(wd.view(np.uint32), wo.view(np.uint32))
    assert not np.array_equal(wd, spec.w)
    assert np.all(wd[plastic == 0] == spec.w[plastic == 0])
    assert_same_state(dev, ora, spec)


def test_stdp_coincident_pre_post_bitexact():
    # groups of neurons with identical state fire in the same steps, so many
    # synapses see their pre and post spike together (LTP then LTD on one
    # synapse, P/src/api/learn/brain.rs:204-286)
    say("stdp coincident")
    spec = random_net(1200, 40, seed=17, model="lif", max_delay=2, noise=0.0)
    v0 = np.repeat(np.linspace(-65.0, -52.0, 6), 200).astype(np.float32)
    spec.pops[0].states[0, :1200] = v0
    spec.w = (np.abs(spec.w) / 200.0).astype(np.float32)
    spec.stdp = dict(a_plus=0.01, a_minus=0.012, tau_plus=20.0, tau_minus=20.0, w_min=0.0, w_max=1.0)
    dev, ora = run_pair(spec, 300)
    assert ora.counters()[0] > 300  # six synchronous groups: ~2300 coincident synapse events
    wd, wo = dev.download_weights(), ora.weights().astype(np.float32)
    assert np.array_equal(wd.view(np.uint32), wo.view(np.uint32))
    assert not np.array_equal(wd, spec.w)
    assert_same_state(dev, ora, spec)


@pytest.mark.parametrize("sched", [SCATTER_TILE, SCATTER_WAVE])
def test_dense_stdp_step_bitexact(sched):
    # CompiledNetwork::stdp_step (pillars-cuda/src/network.rs:495-718): dense rule on every
    # synapse, interleaved with ordinary steps; spikes, weights, state and ring bitwise
    say("dense_stdp", sched)
    spec = random_net(2500, 40, seed=31, model="lif", max_delay=3, sources=6)
    dev = build_device(spec)
    dev.set_scatter_mode(sched)
    ora = po.OracleNetwork(spec, 32)
    args = (0.95, 0.002, 0.0025, -4.0, 12.0)
    for t in range(300):
        if t % 10 < 7:
            dev.stdp_step(*args)
            ora.stdp_step(*args)
        else:
            dev.step()
            ora.step()
        assert np.array_equal(dev.download_spike_ids(), ora.active()), t
    assert ora.counters()[0] > 500
    assert np.array_equal(dev.download_weights().view(np.uint32), ora.weights().astype(np.float32).view(np.uint32))
    assert not np.array_equal(ora.weights().astype(np.float32), spec.w)
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


def _tm_spec(seed, max_delay):
    spec = random_net(3000, 50, seed=seed, model="lif", max_delay=max_delay, sources=12)
    rng = np.random.default_rng(seed)
    n_src = sum(len(s.posts) for s in spec.sources)
    spec.stp = dict(profiles=[(0.5, 100.0, 0.0), (0.1, 1e6, 50.0), (0.3, 20.0, 5.0)], dt=spec.dt,
                    syn_types=rng.integers(0, 4, spec.n_syn).astype(np.uint8),
                    src_types=rng.integers(0, 4, n_src).astype(np.uint8))
    return spec


@pytest.mark.parametrize("max_delay", [0, 7])
def test_short_term_plasticity_bitexact(max_delay):
    # Tsodyks-Markram delivery (push.rs:60-86) on synapses and source edges, mixed profiles,
    # many events per ring cell (the fix-up re-sums delivered, not static, weights)
    say("stp", max_delay)
    spec = _tm_spec(41 + max_delay, max_delay)
    dev, ora = run_pair(spec, 400)
    assert ora.counters()[0] > 500
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)
    # a reset restores R = 1, U = U0 and replays the same trajectory
    dev.reset(spec.pops)
    ora2 = po.OracleNetwork(spec, 32)
    dev.run(150)
    ora2.step(150)
    assert_same_state(dev, ora2, spec)


def test_edge_cases():
    say("edge_cases")
    # n = 1 with a self loop firing every step under strong drive
    spec = single_pop(lif_pop(1, 0.5, 0.5, -70.0, -54.0, -70.0, 0.0, -70.0), 0.5)
    spec.row_ptr = np.array([0, 1], np.uint32)
    spec.col = np.array([0], np.uint32)
    spec.w = np.array([100.0], np.float32)
    spec.delay = np.array([31], np.uint8)
    run_pair(spec, 100, inject=lambda d, o, t: (d.inject_current(0, [0], [200.0]), o.inject_current(0, [0], [200.0])))
    # everyone spikes every step and all land on one post (long chains, overflow path)
    n = 700
    pre = np.arange(n, dtype=np.uint32)
    post = np.zeros(n, np.uint32)
    rp, col, w, d = csr_from_triplets(n, pre, post, np.linspace(-1, 1, n).astype(np.float32), np.zeros(n, np.uint8))
    spec = NetSpec("fanin", [lif_pop(n, 0.5, 0.5, -70.0, -54.0, -70.0, 0.0, -70.0)], rp, col, w, d, 0.5, 0)
    idx = np.arange(n)
    big = np.full(n, 500.0, np.float32)
    dev, ora = run_pair(spec, 20, inject=lambda d, o, t: (d.inject_current(0, idx, big), o.inject_current(0, idx, big)))
    assert_same_ring(dev, ora)
    # mixed models, odd sizes, no synapses at all
    pops = [lif_pop(65, 0.5, 0.1, -70, -54, -70, 2.0, -60.0), izh_pop(1, 0.5, 0.02, 0.2, -65, 8, 0.0, 30, -65, -13),
            hh_pop(130, 0.5, tau_ref=1.0)]
    rp, col, w, d = empty_csr(196)
    spec = NetSpec("mixed", pops, rp, col, w, d, 0.5, 0)
    spec.noise += [(0, np.full(65, 20.0, np.float32), 1), (2, np.full(130, 30.0, np.float32), 2)]
    dev, ora = run_pair(spec, 300)
    assert_same_state(dev, ora, spec)


@pytest.mark.parametrize("sched", [SCATTER_TILE, SCATTER_WAVE])
def test_collisions_across_tiles_and_sources(sched):
    """Many events per ring cell in one step: half of all edges aim at 40 hot
    posts (chains of ~100 per cell, beyond the 8 sorted in LDS), spread over
    79 neuron tiles (several fix-up blocks) and mixed with Poisson sources
    that fire every step into the same cells."""
    say("collisions", sched)
    from simuse.workloads import Source
    rng = np.random.default_rng(5)
    n, deg, hot = 20000, 30, 40
    pre = np.repeat(np.arange(n, dtype=np.uint32), deg)
    post = np.where(rng.random(pre.size) < 0.5, rng.integers(0, hot, pre.size),
                    rng.integers(0, n, pre.size)).astype(np.uint32)
    w = rng.normal(0.0, 1.0, pre.size).astype(np.float32)
    d = rng.integers(1, 4, pre.size).astype(np.uint8)
    rp, col, w, d = csr_from_triplets(n, pre, post, w, d)
    spec = NetSpec("hot", [lif_pop(n, 0.5, 0.5, -70.0, -54.0, -70.0, 0.0, -70.0)], rp, col, w, d, 0.5, 0)
    for s in range(300):
        k = int(rng.integers(1, 6))
        spec.sources.append(Source(1e7, 0.5, 1000 + s, rng.integers(0, hot, k).astype(np.uint32),
                                   rng.normal(0.0, 1.0, k).astype(np.float32),
                                   rng.integers(0, 4, k).astype(np.uint8)))
    kick = [rng.choice(n, 1000, replace=False).astype(np.uint32) for _ in range(25)]
    big = np.full(1000, 500.0, np.float32)
    dev, ora = run_pair(spec, 25, sched=sched,
                        inject=lambda dv, o, t: (dv.inject_current(0, kick[t], big), o.inject_current(0, kick[t], big)))
    assert ora.counters()[1] > 25 * 1000 * deg // 2
    assert dev.counters() == ora.counters()
    assert_same_ring(dev, ora)
    assert_same_state(dev, ora, spec)


def test_reset_and_determinism():
    say("reset_determinism")
    spec = random_net(4000, 80, seed=21, model="izh", max_delay=2)
    dev = build_device(spec)
    dev.record(300)
    dev.run(300)
    r1 = dev.download_raster()
    s1 = dev.download_states(0)
    dev.reset(spec.pops)
    dev.record(300)
    dev.run(300)
    assert np.array_equal(dev.download_raster(), r1)
    assert np.array_equal(dev.download_states(0), s1)
    assert dev.step_counter() == 300


def test_autotune_keeps_state():
    say("autotune")
    spec = random_net(20_000, 100, seed=8, model="izh")
    dev = build_device(spec)
    ora = po.OracleNetwork(spec, 32)
    dev.run(50)
    ora.step(50)
    mode = dev.autotune(8)
    assert mode in (SCATTER_TILE, SCATTER_WAVE)
    dev.run(50)
    ora.step(50)
    assert np.array_equal(dev.download_spike_ids(), ora.active())
    assert dev.step_counter() == 100
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


def test_step_form_switches_and_autotune_bitexact():
    # A network that can take the one-kernel step, switched to the binned form
    # and back mid-run (the pending events move through the ring), then
    # autotuned with both forms among the candidates: spike ids every step,
    # final state, ring and counters equal the uninterrupted oracle.
    from simuse import FORM_BINNED, FORM_ONE_KERNEL
    say("step form switches")
    spec = random_net(20_000, 100, seed=9, model="izh", sources=300)
    dev = build_device(spec)
    ora = po.OracleNetwork(spec, 32)
    forms = []
    for k, on in enumerate([True, False, True, False, None]):
        if on is None:
            dev.autotune(8)
        else:
            dev.set_one_kernel(on)
        forms.append(dev.step_form())
        for t in range(30):
            dev.step() if t % 2 else dev.run(1)
            ora.step()
            assert np.array_equal(dev.download_spike_ids(), ora.active()), (k, t)
        if k == 1:  # a multi-step graph in the binned form too
            dev.run(16)
            ora.step(16)
    assert forms[:4] == [FORM_ONE_KERNEL, FORM_BINNED, FORM_ONE_KERNEL, FORM_BINNED], forms
    assert forms[4] in (FORM_ONE_KERNEL, FORM_BINNED)
    assert dev.counters() == ora.counters() and ora.counters()[0] > 0
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


def test_raster_statistics_match_oracle():
    say("raster_stats")
    spec = random_net(3000, 60, seed=4, model="izh")
    dev = build_device(spec)
    ora = po.OracleNetwork(spec, 32)
    dev.record(1000)
    ora.record(True)
    dev.run(1000)
    ora.step(1000)
    rast = ora.raster()
    rates, cvs = dev.raster_stats(0, 3000, 0.001)
    assert np.allclose(rates, po.firing_rates(rast, 0.001), rtol=1e-6)
    assert np.allclose(cvs, po.isi_cv(rast, 0.001), rtol=1e-4, atol=1e-5)


def test_izhikevich_1m_full_size_bitexact():
    # The benchmark network itself (1M neurons, 100M synapses): 12 single
    # steps with the spike ids of every step, then 88 steps in captured
    # graphs (22 per call), bit for bit on both sides: spike ids and counters
    # after each call, then state and ring.
    say("izhikevich_1m: building")
    spec = izhikevich_network(seed=1, steps=100)
    say("izhikevich_1m: device")
    dev = build_device(spec, graph_steps=11)
    ora = po.OracleNetwork(spec, 32, threads=8)
    for t in range(12):
        dev.step()
        ora.step()
        assert np.array_equal(dev.download_spike_ids(), ora.active()), t
    for c in range(4):
        dev.run(22)
        ora.step(22)
        assert np.array_equal(dev.download_spike_ids(), ora.active()), c
        assert dev.counters() == ora.counters(), c
    assert dev.step_counter() == 100
    assert ora.counters()[0] > 0
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)
    say("izhikevich_1m: ok", ora.counters())


def test_gpea_population_bitwise():
    # gpea.rs pulse protocol (rest, then a constant current) on 1024 neurons with currents
    # from 0 to 3x the 0.5 nA pulse; spikes every step and the final 9 state columns bitwise
    from simuse.models import GPEA_PULSE, gpea_pop
    say("gpea_population")
    n = 1024
    spec = single_pop(gpea_pop(n), 0.025)
    idx = np.arange(n)
    cur = np.linspace(0.0, 3.0 * GPEA_PULSE, n).astype(np.float32)

    def inj(dv, o, t):
        if t >= 400:
            dv.inject_current(0, idx, cur)
            o.inject_current(0, idx, cur)

    dev, ora = run_pair(spec, 2400, inject=inj)
    assert ora.counters()[0] > 1000
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)


def test_gpea_network_bitexact():
    say("gpea_network")
    spec = random_net(3000, 40, seed=61, model="gpea", max_delay=5, noise=60.0, sources=5)
    dev, ora = run_pair(spec, 800, check_every=1)
    assert ora.counters()[0] > 200
    assert dev.counters() == ora.counters()
    assert_same_state(dev, ora, spec)
    assert_same_ring(dev, ora)


@pytest.mark.parametrize("plastic", [False, True])
def test_fixed_out_degree_network_bitexact(plastic):
    # every row holds k edges: the kernels compute the CSR row extents (i k,
    # k) instead of reading row_ptr; one-kernel form, and with pair STDP the
    # two-kernel form whose writers take the extents the same way
    say("fixed out-degree", plastic)
    from simuse.workloads import izhikevich_network
    spec = izhikevich_network(n=30000, k=40, seed=17, steps=300)
    n = sum(p.n for p in spec.pops)
    assert np.array_equal(spec.row_ptr.astype(np.uint64), np.arange(n + 1, dtype=np.uint64) * 40)
    if plastic:
        rng = np.random.default_rng(17)
        spec.w = np.abs(spec.w)
        spec.stdp = dict(a_plus=0.01, a_minus=0.012, tau_plus=20.0, tau_minus=20.0, w_min=0.0, w_max=1.0,
                         plastic=(rng.random(spec.n_syn) < 0.5).astype(np.uint8))
    dev, ora = run_pair(spec, 300, check_every=1, graph_steps=16)
    assert ora.counters()[0] > 1000
    assert_same_state(dev, ora, spec)
    if plastic:
        assert np.array_equal(dev.download_weights().view(np.uint32),
                              ora.weights().astype(np.float32).view(np.uint32))
