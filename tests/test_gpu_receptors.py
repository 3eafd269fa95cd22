"""Conductance-based receptor synapses on the device against the oracle,
bit for bit (needs an MI355X).

The step is the reference's conductance simulation (crates/SiMuse/crates/
pillars-connectome/src/sim/hh.rs:115-166 with pillars/src/kernels/synapse/
kinetics.rs and receptor.rs): decay, external current, receptor currents of
the pre-update potential (NMDA Mg2+ block; 8-lane / tail split at N & ~7),
update, conductance jumps in push order.  The reference's own GPU twin
(pillars-cuda/src/gpea_syn.rs) scatters with float atomics and is checked
only to a tolerance (pillars-connectome/tests/gpea_gpu_syn_parity.rs); here
spike ids every step, state, conductances and ring are compared bitwise.
"""
import sys

import numpy as np
import pytest

import pyoracle as po
from nets import random_net
from simuse import SimuseError
from simuse.models import izh_pop, lif_pop
from simuse.workloads import NetSpec, Source, build_device, csr_from_triplets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def with_receptors(spec, seed, g_scale, current_frac=0.1, ext=None, scale=1.0, codes=None):
    rng = np.random.default_rng(seed)
    n = spec.n_syn
    types = rng.integers(0, 7, n).astype(np.uint8) if codes is None else rng.choice(codes, n).astype(np.uint8)
    types[rng.random(n) < current_frac] = 255           # current-based
    types[rng.random(n) < 0.02] = 9                     # unknown code: AMPA
    w = spec.w.copy()
    cond = types != 255
    w[cond] = np.abs(w[cond]) * g_scale  # conductances are non-negative
    spec.w = w.astype(np.float32)
    spec.receptors = dict(types=types, scale=scale, ext=ext)
    return spec


def run_check(spec, steps, graph_steps=8, tune_at=None):
    dev = build_device(spec, graph_steps=graph_steps)
    ora = po.OracleNetwork(spec, 32)
    total = 0
    for t in range(steps):
        if t == tune_at:
            dev.autotune(8)  # trial steps on a snapshot, conductances included
        dev.step()
        ora.step()
        a, b = dev.download_spike_ids(), ora.active()
        assert np.array_equal(a, b), f"step {t}: spikes differ ({a.size} vs {b.size})"
        total += a.size
    for p, pop in enumerate(spec.pops):
        st = dev.download_states(p)
        for c in range(pop.ns):
            x, y = st[c, : pop.n], ora.state(p, c).astype(np.float32)
            bad = np.nonzero(x.view(np.uint32) != y.view(np.uint32))[0]
            assert bad.size == 0, f"pop {p} col {c}: {bad.size} differ, first {bad[:4]}"
    gd, go = dev.download_conductances(), ora.conductances().astype(np.float32)
    bad = np.argwhere(gd.view(np.uint32) != go.view(np.uint32))
    assert bad.size == 0, f"{len(bad)} conductances differ, first {bad[:4].tolist()}"
    assert np.array_equal(dev.download_ring().view(np.uint32), ora.ring().astype(np.float32).view(np.uint32))
    assert dev.counters() == ora.counters()
    return dev, ora, total


@pytest.mark.parametrize("model,n,steps,g_scale", [
    ("lif", 3001, 300, 0.02),     # N & ~7 = 3000: one tail neuron
    ("izh", 5003, 300, 0.01),
    ("hh", 1201, 400, 0.02),
])
def test_receptor_network_bitexact(model, n, steps, g_scale):
    say("receptors", model, n)
    spec = random_net(n, 25, seed=n, model=model, max_delay=1, sources=3)
    ext = np.zeros(n, np.float32)
    if model == "hh":
        ext[: n // 5] = 15.0  # above the ~6.5 uA/cm2 rheobase
    else:
        ext[: n // 20] = 3.0
    spec = with_receptors(spec, n, g_scale, ext=ext)
    dev, ora, total = run_check(spec, steps, tune_at=40 if model == "lif" else None)
    assert total > 0
    assert np.count_nonzero(dev.download_conductances()) > 0


def test_gpea_syn_reference_setup_bitexact():
    # pillars-connectome/tests/gpea_gpu_syn_parity.rs: 1024 GPEA neurons, 12
    # synapses each, AMPA / nAChR / GABA_A / GluCl, tonic 12 on 16 neurons, 50 ms
    say("gpea syn")
    n, deg = 1024, 12
    rng = np.random.default_rng(0xFEEDBEEF)
    pre = np.repeat(np.arange(n, dtype=np.uint32), deg)
    post = rng.integers(0, n, pre.size).astype(np.uint32)
    w = (0.35 * (0.5 + rng.integers(0, 1000, pre.size) / 1000.0)).astype(np.float32)
    types = rng.choice(np.array([0, 0, 0, 4, 4, 4, 2, 2, 2, 5], np.uint8), pre.size)
    rp, col, ww, d = csr_from_triplets(n, pre, post, w, np.zeros(pre.size, np.uint8))
    # csr_from_triplets sorts stably by pre, which the repeat already is
    spec = random_net(n, 1, seed=1, model="gpea", noise=0.0)
    spec = NetSpec("gpea_syn", spec.pops, rp, col, ww, d, spec.dt, 2000)
    ext = np.zeros(n, np.float32)
    ext[:16] = 12.0
    spec.receptors = dict(types=types, scale=1.0, ext=ext)
    dev, ora, total = run_check(spec, 2000, graph_steps=16)
    assert total > 0


def test_receptors_mixed_populations_and_sources():
    # two models (the two-kernel binned step with several populations), the
    # lane/tail split falling inside the second population, noise, sources
    say("receptors mixed")
    rng = np.random.default_rng(4)
    n1, n2 = 1000, 1507
    lif = lif_pop(n1, 0.1, 0.05, -48.0, -50.0, -65.0, 2.0, rng.uniform(-65, -50, n1).astype(np.float32))
    v0 = np.full(n2, -65.0, np.float32)
    izh = izh_pop(n2, 0.1, 0.02, 0.2, -65.0, 8.0, 0.0, 30.0, v0, (0.2 * v0).astype(np.float32))
    N = n1 + n2
    deg = rng.poisson(20, N)
    pre = np.repeat(np.arange(N, dtype=np.uint32), deg)
    post = rng.integers(0, N, pre.size).astype(np.uint32)
    w = (rng.random(pre.size) * 0.05).astype(np.float32)
    d = rng.integers(0, 2, pre.size).astype(np.uint8)
    rp, col, w, d = csr_from_triplets(N, pre, post, w, d)
    spec = NetSpec("mixed_rec", [lif, izh], rp, col, w, d, 0.1, 300)
    spec.noise = [(1, np.full(n2, 4.0, np.float32), 77)]
    spec.sources.append(Source(2000.0, 0.1, 99, rng.integers(0, N, 300).astype(np.uint32),
                               (rng.random(300) * 10).astype(np.float32), np.ones(300, np.uint8)))
    spec = with_receptors(spec, 5, 1.0, scale=0.5)
    _, _, total = run_check(spec, 300)
    assert total > 0


def test_receptors_with_pair_stdp_and_tm():
    say("receptors + stdp + tm")
    spec = random_net(2500, 30, seed=8, model="lif", max_delay=1, sources=4)
    spec = with_receptors(spec, 8, 0.02)
    rng = np.random.default_rng(8)
    spec.w = np.abs(spec.w)
    spec.stdp = dict(a_plus=0.01, a_minus=0.012, tau_plus=20.0, tau_minus=20.0, w_min=0.0, w_max=1.0,
                     plastic=(rng.random(spec.n_syn) < 0.5).astype(np.uint8))
    dev, ora, total = run_check(spec, 300)
    assert np.array_equal(dev.download_weights().view(np.uint32), ora.weights().astype(np.float32).view(np.uint32))
    spec2 = random_net(2500, 30, seed=9, model="lif", max_delay=1, sources=4)
    spec2 = with_receptors(spec2, 9, 0.02)
    n_src = sum(len(s.posts) for s in spec2.sources)
    spec2.stp = dict(profiles=[(0.5, 100.0, 0.0), (0.1, 1e6, 50.0)], dt=spec2.dt,
                     syn_types=rng.integers(0, 3, spec2.n_syn).astype(np.uint8),
                     src_types=rng.integers(0, 3, n_src).astype(np.uint8))
    run_check(spec2, 300)


def test_receptor_hub_posts_long_lists():
    # every neuron also projects onto 4 hub posts: hundreds of events per hub
    # per step (the consumer's sorted long lists, receptor codes in the keys)
    say("receptor hubs")
    n = 3000
    spec = random_net(n, 10, seed=12, model="lif", max_delay=1, sources=0)
    rng = np.random.default_rng(12)
    m = spec.n_syn
    pre = np.repeat(np.arange(n, dtype=np.uint32), np.diff(spec.row_ptr).astype(np.int64))
    hub_pre = np.repeat(np.arange(n, dtype=np.uint32), 4)
    hub_post = np.tile(np.arange(4, dtype=np.uint32), n)
    pre2 = np.concatenate([pre, hub_pre])
    post2 = np.concatenate([spec.col, hub_post])
    w2 = np.concatenate([spec.w, (rng.random(hub_pre.size) * 0.01).astype(np.float32)])
    d2 = np.concatenate([spec.delay, np.ones(hub_pre.size, np.uint8)])
    order = np.argsort(pre2, kind="stable")
    rp, col, w, d = csr_from_triplets(n, pre2[order], post2[order], w2[order], d2[order])
    spec = NetSpec("hubs", spec.pops, rp, col, w, d, spec.dt, 200, noise=spec.noise)
    # a strong external current keeps ~30 neurons firing per step
    spec = with_receptors(spec, 13, 0.02, ext=np.full(n, 30.0, np.float32))
    assert spec.n_syn == m + 4 * n
    _, _, total = run_check(spec, 200)
    assert total > 2000


def test_receptor_reset_and_long_delay_refused():
    say("receptor reset / refusal")
    spec = random_net(1500, 20, seed=21, model="lif", max_delay=1, sources=2)
    spec = with_receptors(spec, 21, 0.02)
    dev = build_device(spec)
    dev.run(50)
    assert np.count_nonzero(dev.download_conductances()) > 0
    dev.reset(spec.pops)
    assert not np.any(dev.download_conductances())
    ora = po.OracleNetwork(spec, 32)
    dev.run(80)
    ora.step(80)
    assert np.array_equal(dev.download_conductances().view(np.uint32),
                          ora.conductances().astype(np.float32).view(np.uint32))
    # the chained scatter (a delay > 1) has no receptor form
    bad = with_receptors(random_net(500, 10, seed=22, model="lif", max_delay=3), 22, 0.02)
    net = build_device(bad)
    with pytest.raises(SimuseError, match="receptor synapses need"):
        net.step()
    # wrong sizes are refused on the host
    net2 = build_device(random_net(500, 10, seed=23, model="lif", max_delay=1))
    with pytest.raises(ValueError):
        net2.set_receptors(np.zeros(3, np.uint8))
    with pytest.raises(SimuseError, match="scale must be finite"):
        net2.set_receptors(np.zeros(net2.n_syn, np.uint8), float("nan"))


def test_receptor_types_without_synapses(tmp_path):
    # only AMPA and GABA_A synapses: the other five conductances stay zero and
    # the step skips their columns; a snapshot whose file holds a nonzero
    # conductance of such a type (here GABA_B and NMDA) brings them back
    say("dead receptor types")
    from simuse.formats import TaasReader, TaasWriter
    n = 2000
    spec = random_net(n, 20, seed=31, model="lif", max_delay=1, sources=2)
    ext = np.zeros(n, np.float32)
    ext[: n // 10] = 3.0
    spec = with_receptors(spec, 31, 0.02, ext=ext, codes=[0, 2])
    dev, ora, total = run_check(spec, 120)
    assert total > 0
    g = dev.download_conductances()
    assert np.count_nonzero(g[[0, 2]]) > 0 and not np.any(g[[1, 3, 4, 5, 6]])
    path = tmp_path / "s.taas"
    dev.save_state(path)
    r = TaasReader.from_file(path)
    w = TaasWriter()
    for k, v in r.metadata().items():
        w.add_metadata(k, v)
    rng = np.random.default_rng(31)
    for s in r.list_sections():
        raw = r._read(s)
        if s.name == "receptors.g":
            gg = np.frombuffer(raw, np.float32).reshape(7, n).copy()
            gg[3, rng.integers(0, n, 300)] = 0.5   # GABA_B
            gg[1, rng.integers(0, n, 300)] = 0.25  # NMDA
            raw = gg.tobytes()
        w.push(s.name, s.kind, s.count, raw)
    edited = tmp_path / "edited.taas"
    w.to_file(edited)
    dev2 = build_device(spec)
    dev2.load_state(edited)
    ora2 = po.OracleNetwork(spec, 32)
    ora2.set_state_from_snapshot(edited)
    for t in range(150):
        dev2.step()
        ora2.step()
        assert np.array_equal(dev2.download_spike_ids(), ora2.active()), t
    gd, go = dev2.download_conductances(), ora2.conductances().astype(np.float32)
    assert np.count_nonzero(gd[3]) > 0 and np.count_nonzero(gd[1]) > 0
    assert np.array_equal(gd.view(np.uint32), go.view(np.uint32))
    st = dev2.download_states(0)
    assert np.array_equal(st[0, :n].view(np.uint32), ora2.state(0, 0).astype(np.float32).view(np.uint32))
