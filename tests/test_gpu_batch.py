"""Batched networks: same results alone and in a mix; device-generated
networks against the oracle; per-network horizons (needs an MI355X)."""
import numpy as np
import pytest

import pyoracle as po
from nets import random_net
from simuse import MODEL_HH, MODEL_IZHIKEVICH, MODEL_LIF
from simuse.batch import Batch, gen_desc
from simuse.workloads import NetSpec, Source, build_device, spec_from_generated

pytestmark = pytest.mark.gpu

@pytest.mark.parametrize("model", [MODEL_IZHIKEVICH, MODEL_LIF, MODEL_HH])
def test_generated_network_matches_oracle(model):
    print("generated", model, flush=True)
    if model == MODEL_IZHIKEVICH:
        g = gen_desc(model, 3000, 60, seed=11, dt=1.0, w_exc=0.5, w_inh=1.0, drive_rate_hz=40.0,
                     drive_weight=8.0, noise_sigma=3.0)
    elif model == MODEL_HH:
        g = gen_desc(model, 2000, 40, seed=13, dt=0.01, w_exc=2.0, w_inh=6.0, d_min=1, d_max=4,
                     drive_rate_hz=2000.0, drive_weight=40.0, noise_sigma=30.0)
    else:
        g = gen_desc(model, 3000, 60, seed=12, dt=0.1, w_exc=20.0, w_inh=60.0, d_min=1, d_max=15,
                     drive_rate_hz=3000.0, drive_weight=60.0, noise_sigma=0.0)
    T = 1500 if model == MODEL_HH else 300
    b = Batch()
    nid = b.add_generated(g)
    spec = spec_from_generated(b, nid, g, T)
    o = po.OracleNetwork(spec, 32)
    b.record(nid, T)
    o.record(True)
    b.run(T)
    o.step(T)
    bits = b.download_raster_bits(nid)
    ras = o.raster()
    assert ras.sum() > 100
    # compare the device bit raster with the oracle raster population by population
    off, w0 = 0, 0
    for pop in spec.pops:
        nt = (pop.n + 255) // 256
        un = np.unpackbits(bits[:, w0: w0 + 4 * nt].copy().view(np.uint8).reshape(bits.shape[0], -1),
                           axis=1, bitorder="little")[:, : pop.n]
        assert np.array_equal(un, ras[:, off: off + pop.n])
        off += pop.n
        w0 += 4 * nt
    for p, pop in enumerate(spec.pops):
        st = b.download_state(nid, p)
        for c in range(pop.ns):
            assert np.array_equal(st[c, : pop.n].view(np.uint32),
                                  o.state(p, c).astype(np.float32).view(np.uint32))
    assert b.counters(nid) == o.counters()


def test_drive_equals_explicit_sources():
    print("drive", flush=True)
    spec = random_net(2000, 40, seed=9, model="lif", max_delay=5, noise=0.0)
    a = build_device(spec)
    a.set_poisson_drive(0, 2500.0, spec.dt, 80.0, 777)
    spec2 = random_net(2000, 40, seed=9, model="lif", max_delay=5, noise=0.0)
    for i in range(2000):
        spec2.sources.append(Source(2500.0, spec.dt, (777 + i * 2654435761) & 0xFFFFFFFF,
                                    np.array([i], np.uint32), np.array([80.0], np.float32), np.array([0], np.uint8)))
    bdev = build_device(spec2)
    o = po.OracleNetwork(spec2, 32)
    for t in range(300):
        a.step()
        bdev.step()
        o.step()
        ids = o.active()
        assert np.array_equal(a.download_spike_ids(), ids), t
        assert np.array_equal(bdev.download_spike_ids(), ids), t
    assert o.counters()[0] > 100


def test_batch_alone_vs_mix_bitexact():
    print("alone_vs_mix", flush=True)
    specs = [random_net(1500, 50, seed=1, model="izh", max_delay=3),
             random_net(700, 30, seed=2, model="lif", max_delay=15, sources=4),
             random_net(3000, 80, seed=3, model="izh", max_delay=0),
             random_net(500, 20, seed=4, model="hh", max_delay=2, noise=40.0),
             random_net(64, 10, seed=5, model="lif", max_delay=31)]
    T = 400
    alone = []
    for s in specs:
        d = build_device(s)
        d.record(T)
        d.run(T)
        alone.append((d.download_raster(), [d.download_states(p) for p in range(len(s.pops))], d.counters()))
        d.close()
    b = Batch()
    ids = [b.add_spec(s) for s in specs]
    g = gen_desc(MODEL_IZHIKEVICH, 5000, 100, seed=77, dt=1.0, drive_rate_hz=30.0, drive_weight=6.0,
                 noise_sigma=2.0)
    b.add_generated(g)  # an extra network in the mix
    for i in ids:
        b.record(i, T)
    b.set_graph_steps(7)
    b.run(T)
    for s, i, (ras, sts, cnt) in zip(specs, ids, alone):
        bits = b.download_raster_bits(i)
        un = np.unpackbits(bits.copy().view(np.uint8).reshape(bits.shape[0], -1), axis=1, bitorder="little")
        # single-population specs: the first n bits are the raster
        n = s.pops[0].n
        assert np.array_equal(un[:, :n], ras[:, :n])
        for p in range(len(s.pops)):
            assert np.array_equal(b.download_state(i, p).view(np.uint32), sts[p].view(np.uint32))
        assert b.counters(i) == cnt


def test_horizon_and_removal():
    print("horizon", flush=True)
    s1 = random_net(2000, 50, seed=21, model="izh")
    s2 = random_net(2000, 50, seed=22, model="izh")
    ref = build_device(s1)
    ref.run(50)
    st50 = ref.download_states(0)
    ref2 = build_device(s2)
    ref2.run(120)
    st120 = ref2.download_states(0)
    b = Batch()
    i1 = b.add_spec(s1)
    i2 = b.add_spec(s2)
    b.set_horizon(i1, 50)
    b.run(100)
    assert b.step_counter(i1) == 50 and b.step_counter(i2) == 100
    assert np.array_equal(b.download_state(i1, 0), st50)
    b.remove(i1)
    assert b.n_networks() == 1
    b.run(20)
    assert np.array_equal(b.download_state(i2, 0), st120)


def test_batch_with_plasticity_alone_vs_mix():
    # pair STDP and Tsodyks-Markram networks keep their results inside a batch
    print("plastic_mix", flush=True)
    rng = np.random.default_rng(3)
    a = random_net(2000, 40, seed=51, model="lif", max_delay=4, sources=5)
    a.w = np.abs(a.w) / 40.0
    a.stdp = dict(a_plus=0.01, a_minus=0.012, tau_plus=20.0, tau_minus=20.0, w_min=0.0, w_max=1.0)
    b = random_net(2500, 40, seed=52, model="lif", max_delay=0, sources=5)
    n_src = sum(len(s.posts) for s in b.sources)
    b.stp = dict(profiles=[(0.5, 100.0, 0.0), (0.2, 50.0, 20.0)], dt=b.dt,
                 syn_types=rng.integers(0, 3, b.n_syn).astype(np.uint8),
                 src_types=rng.integers(0, 3, n_src).astype(np.uint8))
    c = random_net(1500, 30, seed=53, model="izh", max_delay=2)
    specs = [a, b, c]
    T = 300
    alone = []
    for s in specs:
        d = build_device(s)
        d.run(T)
        alone.append(([d.download_states(p) for p in range(len(s.pops))], d.counters()))
        d.close()
    bt = Batch()
    ids = [bt.add_spec(s) for s in specs]
    bt.set_graph_steps(5)
    bt.run(T)
    for s, i, (sts, cnt) in zip(specs, ids, alone):
        for p in range(len(s.pops)):
            assert np.array_equal(bt.download_state(i, p).view(np.uint32), sts[p].view(np.uint32))
        assert bt.counters(i) == cnt
