"""Checkpoint / resume of a network's device state through `.taas` v2 files
(needs an MI355X).

A network runs k1 steps and saves; a second network built the same way (and
stepped elsewhere first, so that nothing of its own state survives by luck)
loads the file and both step on: spikes of every step, the final FP32 state,
weights, ring and counters must agree bit for bit, and the resumed run must
equal the oracle's uninterrupted run.  Every scatter form is covered: the
one-kernel binned step, the two-kernel binned step (mixed models), the
chained form (delays up to 12), pair STDP, the dense stdp_step rule and
Tsodyks-Markram synapses.  Reference: crates/SiMuse/crates/pillars-cuda/src/
network.rs:481,721-748,827-842 (the state a resume restores) and
taas-format/src/lib.rs:8-24 (the container).
"""
import sys

import numpy as np
import pytest

import pyoracle as po
from nets import random_net, two_pop_net
from simuse import SimuseError
from simuse.formats import ConnectomeData, TaasReader, TaasWriter
from simuse.models import lif_pop
from simuse.workloads import NetSpec, build_device

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _spec(kind):
    if kind == "binned_izh":
        return random_net(6000, 60, seed=3, model="izh", max_delay=1, sources=4)
    if kind == "binned_mixed":
        spec = two_pop_net(3000)  # LIF -> Izhikevich (delay 1) and back (delay 0), driven by noise
        spec.noise = [(0, np.full(3000, 20.0, np.float32), 5), (1, np.full(3000, 20.0, np.float32), 6)]
        return spec
    if kind == "chained_lif":
        return random_net(3000, 50, seed=5, model="lif", max_delay=12, sources=8)
    if kind == "hh":
        return random_net(1500, 30, seed=7, model="hh", max_delay=2, sources=3)
    if kind == "stdp":
        spec = random_net(3000, 50, seed=13, model="lif", max_delay=4, sources=10)
        rng = np.random.default_rng(0)
        spec.w = np.abs(spec.w) / 40.0
        spec.stdp = dict(a_plus=0.01, a_minus=0.012, tau_plus=20.0, tau_minus=20.0, w_min=0.0, w_max=1.0,
                         plastic=(rng.random(spec.n_syn) < 0.7).astype(np.uint8))
        return spec
    if kind == "tm":
        spec = random_net(3000, 50, seed=48, model="lif", max_delay=7, sources=12)
        rng = np.random.default_rng(48)
        n_src = sum(len(s.posts) for s in spec.sources)
        spec.stp = dict(profiles=[(0.5, 100.0, 0.0), (0.1, 1e6, 50.0), (0.3, 20.0, 5.0)], dt=spec.dt,
                        syn_types=rng.integers(0, 4, spec.n_syn).astype(np.uint8),
                        src_types=rng.integers(0, 4, n_src).astype(np.uint8))
        return spec
    if kind == "dense_stdp":
        return random_net(2500, 40, seed=31, model="lif", max_delay=3, sources=6)
    if kind == "receptors":
        spec = random_net(3001, 30, seed=61, model="lif", max_delay=1, sources=4)
        rng = np.random.default_rng(61)
        types = rng.integers(0, 7, spec.n_syn).astype(np.uint8)
        types[rng.random(spec.n_syn) < 0.1] = 255
        spec.w = np.where(types != 255, np.abs(spec.w) * 0.02, spec.w).astype(np.float32)
        ext = np.zeros(3001, np.float32)
        ext[:100] = 2.0
        spec.receptors = dict(types=types, scale=1.0, ext=ext)
        return spec
    raise ValueError(kind)


DENSE = (0.95, 0.002, 0.0025, -4.0, 12.0)


def _dense(kind, t):
    """The dense rule runs on 7 of every 10 steps after the first 40."""
    return kind == "dense_stdp" and t >= 40 and t % 10 < 7


def _advance(net, kind, t0, n, ids=None):
    for t in range(t0, t0 + n):
        if _dense(kind, t):
            net.stdp_step(*DENSE)
        else:
            net.step()
        if ids is not None:
            ids.append(net.download_spike_ids())


def _oracle_run(spec, kind, n):
    ora = po.OracleNetwork(spec, 32)
    for t in range(n):
        if _dense(kind, t):
            ora.stdp_step(*DENSE)
        else:
            ora.step()
    return ora


def _assert_same(a, b, spec):
    for p, pop in enumerate(spec.pops):
        sa, sb = a.download_states(p), b.download_states(p)
        assert np.array_equal(sa[:, : pop.n].view(np.uint32), sb[:, : pop.n].view(np.uint32)), f"pop {p} state"
    assert np.array_equal(a.download_weights().view(np.uint32), b.download_weights().view(np.uint32)), "weights"
    assert np.array_equal(a.download_ring().view(np.uint32), b.download_ring().view(np.uint32)), "ring"
    assert a.counters() == b.counters()
    assert a.step_counter() == b.step_counter()


@pytest.mark.parametrize("kind", ["binned_izh", "binned_mixed", "chained_lif", "hh", "stdp", "tm", "dense_stdp",
                                  "receptors"])
def test_checkpoint_resume_bitexact(kind, tmp_path):
    say("checkpoint", kind)
    spec = _spec(kind)
    k1, k2 = 121, 100  # an odd step: the binned records alternate buffers by step parity
    path = str(tmp_path / "a.taas")
    a = build_device(spec)
    a.step_graph(40)  # graph-launched steps before the save
    a.sync()
    _advance(a, kind, 40, k1 - 40)
    a.save_state(path)
    ids_a = []
    _advance(a, kind, k1, k2, ids_a)

    b = build_device(spec)
    _advance(b, kind, 0, 37)  # a different state and step counter
    b.load_state(path)
    assert b.step_counter() == k1
    # what load restored is everything save sees
    path_b = str(tmp_path / "b.taas")
    b.save_state(path_b)
    ra, rb = TaasReader.from_file(path), TaasReader.from_file(path_b)
    ra.validate()
    assert [s.name for s in ra.list_sections()] == [s.name for s in rb.list_sections()]
    for s in ra.list_sections():
        assert ra.get_blob(s.name) == rb.get_blob(s.name), s.name
    ids_b = []
    _advance(b, kind, k1, k2, ids_b)
    for t, (x, y) in enumerate(zip(ids_a, ids_b)):
        assert np.array_equal(x, y), f"step {k1 + t}: spikes differ ({x.size} vs {y.size})"
    if kind != "hh":
        assert sum(x.size for x in ids_a) > 0
    _assert_same(a, b, spec)

    # the resumed trajectory is the reference's uninterrupted one
    ora = _oracle_run(spec, kind, k1 + k2)
    assert np.array_equal(b.download_spike_ids(), ora.active())
    for p, pop in enumerate(spec.pops):
        st = b.download_states(p)
        for c in range(pop.ns):
            assert np.array_equal(st[c, : pop.n].view(np.uint32), ora.state(p, c).astype(np.float32).view(np.uint32))
    assert np.array_equal(b.download_weights().view(np.uint32), ora.weights().astype(np.float32).view(np.uint32))
    assert np.array_equal(b.download_ring().view(np.uint32), ora.ring().astype(np.float32).view(np.uint32))
    if spec.receptors:
        assert np.array_equal(b.download_conductances().view(np.uint32),
                              ora.conductances().astype(np.float32).view(np.uint32))


def test_snapshot_file_contents(tmp_path):
    say("snapshot contents")
    spec = random_net(2000, 30, seed=11, model="izh", max_delay=0, sources=2)
    a = build_device(spec)
    a.run(25)
    path = tmp_path / "s.taas"
    a.save_state(path)
    r = TaasReader.from_file(path)
    r.validate()
    assert r.manifest.game == "simuse"
    assert r.metadata()["content"] == "network-state"
    assert r.metadata()["step"] == "25"
    assert r.get_u64("step").tolist() == [25]
    st = a.download_states(0)
    pop = spec.pops[0]
    assert np.array_equal(r.get_f32("pop0.states").reshape(pop.ns, pop.n), st[:, : pop.n])
    assert np.array_equal(r.get_f32("ring").reshape(32, -1), a.download_ring())
    assert np.array_equal(r.get_f32("weights"), a.download_weights())
    ids = a.download_spike_ids()
    cnt = r.get_i32("spikes.count")
    assert int(cnt.sum()) == ids.size
    assert a.counters() == (int(r.get_u64("counters.spikes").sum()), int(r.get_u64("counters.events").sum()))


def test_load_rejects_and_leaves_network_unchanged(tmp_path):
    say("load errors")
    spec = random_net(2000, 30, seed=21, model="lif", max_delay=5, sources=2)
    a = build_device(spec)
    a.run(30)
    path = tmp_path / "a.taas"
    a.save_state(path)

    # a different CSR (one target moved)
    spec2 = random_net(2000, 30, seed=21, model="lif", max_delay=5, sources=2)
    spec2.col = spec2.col.copy()
    spec2.col[7] = (spec2.col[7] + 1) % spec2.n_total
    b = build_device(spec2)
    b.run(10)
    before = b.download_states(0).copy(), b.download_ring().copy(), b.step_counter()
    with pytest.raises(SimuseError, match="CSR col"):
        b.load_state(path)
    assert np.array_equal(b.download_states(0), before[0])
    assert np.array_equal(b.download_ring(), before[1])
    assert b.step_counter() == before[2]

    # another noise seed
    spec3 = random_net(2000, 30, seed=21, model="lif", max_delay=5, sources=2)
    spec3.noise = [(0, spec3.noise[0][1], spec3.noise[0][2] + 1)]
    with pytest.raises(SimuseError, match="noise"):
        build_device(spec3).load_state(path)

    # corrupted payload, truncated file, not a snapshot, missing file
    c = build_device(spec)
    c.run(5)
    raw = bytearray(path.read_bytes())
    raw[-3] ^= 0x55
    bad = tmp_path / "bad.taas"
    bad.write_bytes(bytes(raw))
    with pytest.raises(SimuseError, match="CRC mismatch"):
        c.load_state(bad)
    bad.write_bytes(bytes(raw[: len(raw) // 2]))
    with pytest.raises(SimuseError, match="exceeds"):
        c.load_state(bad)
    TaasWriter().f32("x", [1.0]).to_file(bad)
    with pytest.raises(SimuseError, match="not a network state"):
        c.load_state(bad)
    with pytest.raises(SimuseError, match="cannot open"):
        c.load_state(tmp_path / "missing.taas")
    assert c.step_counter() == 5
    # the good file still loads and matches the saver
    c.load_state(path)
    _assert_same(a, c, spec)


def test_connectome_baked_runs_like_oracle(tmp_path):
    # a baked ConnectomeData file wired as Brain::from_connectome
    # (pillars-connectome/src/brain.rs:19-29,56-83) steps like the oracle on its CSR
    say("connectome")
    rng = np.random.default_rng(5)
    n = 2500
    deg = rng.poisson(40, n)
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    col = rng.integers(0, n, rp[-1]).astype(np.int32)
    nt = np.where(rng.random(n) < 0.8, 1.0, -1.0).astype(np.float32)
    pre = np.repeat(np.arange(n), deg)
    vals = (rng.random(rp[-1]) * np.where(nt[pre] > 0, 8.0, -25.0)).astype(np.float32)
    conn = ConnectomeData(rp, col, vals, nt)
    conn.build_receptor_types()
    p = tmp_path / "c.bin"
    conn.save_baked(p)
    loaded = ConnectomeData.load_baked(p)
    pop = lif_pop(n, 0.1, 0.05, -48.0, -50.0, -65.0, 2.0, rng.uniform(-65, -50, n).astype(np.float32))
    net = loaded.to_network(pop)
    spec = NetSpec("connectome", [pop], *loaded.csr(), 0.1, 200)
    ora = po.OracleNetwork(spec, 32)
    total = 0
    for t in range(200):
        net.step()
        ora.step()
        ids = net.download_spike_ids()
        assert np.array_equal(ids, ora.active()), t
        total += ids.size
    assert total > 200


@pytest.mark.parametrize("kind", ["stdp", "receptors", "binned_izh"])
def test_device_snapshot_resumes_in_oracle(kind, tmp_path):
    # the snapshot holds everything the step depends on: the FP32 oracle twin
    # loaded from the device's file continues exactly as the device does
    say("snapshot into oracle", kind)
    spec = _spec(kind)
    dev = build_device(spec)
    dev.run(90)
    path = tmp_path / "d.taas"
    dev.save_state(path)
    ora = po.OracleNetwork(spec, 32)
    ora.set_state_from_snapshot(path)
    for t in range(100):
        dev.step()
        ora.step()
        assert np.array_equal(dev.download_spike_ids(), ora.active()), t
    for p, pop in enumerate(spec.pops):
        st = dev.download_states(p)
        for c in range(pop.ns):
            assert np.array_equal(st[c, : pop.n].view(np.uint32), ora.state(p, c).astype(np.float32).view(np.uint32))
    assert np.array_equal(dev.download_weights().view(np.uint32), ora.weights().astype(np.float32).view(np.uint32))
    if spec.receptors:
        assert np.array_equal(dev.download_conductances().view(np.uint32),
                              ora.conductances().astype(np.float32).view(np.uint32))
