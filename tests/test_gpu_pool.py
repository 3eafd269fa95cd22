"""The creature pool on the MI355X (per-post pull scatter, alpha-LIF).

* the HIP pool against the oracle pool bit for bit: spike history of every
  substep and the final state, with per-creature inputs, weight syncs between
  ticks, a creature registered mid-run (which restarts the pool), submits
  longer than 64 substeps, and per-synapse delays;
* the HIP pool against the HIP engine's push scatter on the same network
  (alpha-LIF population, signed weights, delay 1 + max(d, 1), the tick input
  applied with inject_voltage) -- two independent device algorithms;
* the boundary's error paths.
"""
import numpy as np
import pytest

import pyoracle as po
from simuse import SimuseError
from simuse.pool import BrainPool
from simuse.workloads import build_device, creature, pool_net_spec

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _register(pools, c):
    for p in pools:
        p.register(c.row_ptr, c.col, c.values, c.nt)


def _same_state(dev, ora):
    for col in range(3):
        assert np.array_equal(_bits(dev.download_state(col)), _bits(ora.state(col))), col


def test_pool_matches_oracle_with_syncs_registration_and_long_submits():
    print("pool vs oracle: 12 creatures, syncs, mid-run registration", flush=True)
    rng = np.random.default_rng(11)
    sizes = [int(x) for x in rng.integers(8, 400, 12)]
    cs = [creature(n, p=float(rng.uniform(0.03, 0.15)), seed=100 + k) for k, n in enumerate(sizes)]
    dev, ora = BrainPool(), po.OraclePool()
    for c in cs[:9]:
        _register((dev, ora), c)
    live = cs[:9]
    total = 0
    for t in range(24):
        if t == 10:  # three more creatures: the pool is rebuilt, everyone restarts
            for c in cs[9:]:
                _register((dev, ora), c)
            live = cs
        if t % 4 == 3:  # weight sync of some creatures
            mirrors = []
            for k, c in enumerate(live):
                if k % 2 == t % 2:
                    w = (c.values * rng.uniform(0.7, 1.3, c.n_syn)).astype(np.float32)
                    mirrors.append((w, True))
                    ora.set_weights(k, w)
                else:
                    mirrors.append((None, False))
            dev.sync_weights(mirrors)
        exts = [rng.uniform(0.0, 9.0, c.n).astype(np.float32) if (t + k) % 3 else None for k, c in enumerate(live)]
        spt = 130 if t in (5, 17) else 50
        flat, slots = dev.step(exts, spt)
        ref = ora.step(exts, spt)
        assert flat.shape == ref.shape and np.array_equal(flat, ref), t
        assert [s.n for s in slots] == [c.n for c in live]
        total += int(flat.sum())
        if t in (9, 10, 23):
            _same_state(dev, ora)
    n, syn, steps, delay = dev.info()
    print(f"  {n} neurons, {syn} synapses, {total} spikes, delay {delay}", flush=True)
    assert total > 5000 and delay == 5 and steps == 50 * 13 + 130  # ticks 10..23 since the rebuild
    dev.close()


def test_pool_per_synapse_delays_and_overrides_match_oracle():
    print("pool vs oracle: per-synapse delays, delay overrides", flush=True)
    rng = np.random.default_rng(3)
    cs = [creature(n, p=0.1, seed=200 + n) for n in (90, 300, 41)]
    for delay in (0, 1, 12, 31):
        dev, ora = BrainPool(delay_steps=delay), po.OraclePool(delay_steps=delay)
        for c in cs:
            _register((dev, ora), c)
        if delay == 12:
            for k, c in enumerate(cs):
                d = rng.integers(0, 40, c.n_syn).astype(np.uint8)
                dev.set_delays(k, d)
                ora.set_delays(k, d)
        total = 0
        for t in range(8):
            exts = [rng.uniform(0.0, 9.0, c.n).astype(np.float32) for c in cs]
            flat, _ = dev.step(exts, 70)
            assert np.array_equal(flat, ora.step(exts, 70)), (delay, t)
            total += int(flat.sum())
        _same_state(dev, ora)
        assert total > 300, delay
        dev.close()


def test_pool_creature_sizes_at_and_above_one_workgroup_match_oracle():
    """Creatures of at most 1024 neurons run a submit chunk as one launch (one
    block per creature, substeps separated by block barriers); a creature
    above 1024 sends the pool to one launch per substep.  Both against the
    oracle, bit for bit, with in-degrees above the 32 in-edges a neuron holds
    in registers."""
    rng = np.random.default_rng(5)
    for sizes in ((1024, 64, 3), (1500, 230, 700)):
        print(f"pool vs oracle: creatures {sizes}", flush=True)
        cs = [creature(n, p=0.04, seed=400 + k) for k, n in enumerate(sizes)]
        dev, ora = BrainPool(), po.OraclePool()
        for c in cs:
            _register((dev, ora), c)
        total = 0
        for t in range(6):
            exts = [rng.uniform(0.0, 9.0, c.n).astype(np.float32) for c in cs]
            flat, _ = dev.step(exts, 70)
            assert np.array_equal(flat, ora.step(exts, 70)), (sizes, t)
            total += int(flat.sum())
        _same_state(dev, ora)
        assert total > 1000, sizes
        dev.close()


def test_pool_equals_engine_push_scatter():
    """The engine's deterministic push scatter (chained form: delay 6) and the
    pool's pull scatter give the same spikes and state on the same pool."""
    print("pool vs engine push scatter: 96 creatures", flush=True)
    rng = np.random.default_rng(21)
    cs = [creature(int(n), p=0.06, seed=300 + k) for k, n in enumerate(rng.integers(50, 700, 96))]
    dev = BrainPool()
    for c in cs:
        dev.register(c.row_ptr, c.col, c.values, c.nt)
    spec = pool_net_spec(cs, 5)
    net = build_device(spec, graph_steps=25)
    n = spec.n_total
    total = 0
    for t in range(16):
        exts = [rng.uniform(0.0, 9.0, c.n).astype(np.float32) if (t + k) % 4 else None for k, c in enumerate(cs)]
        flat, _ = dev.step(exts, 50)
        full = np.concatenate([e if e is not None else np.zeros(c.n, np.float32) for e, c in zip(exts, cs)])
        net.inject_voltage(0, None, full)
        net.record(50)
        net.run(50)
        raster = net.download_raster()
        assert raster.shape == (50, n)
        assert np.array_equal(flat.reshape(50, n), raster.astype(np.uint32)), t
        total += int(flat.sum())
    st = net.download_states(0)
    for col in range(3):
        assert np.array_equal(_bits(dev.download_state(col)), _bits(st[col, :n])), col
    print(f"  {n} neurons, {spec.n_syn} synapses, {total} spikes in 800 substeps, form {net.step_form()}", flush=True)
    assert total > 20_000
    net.close()
    dev.close()


def test_pool_boundary_errors():
    dev = BrainPool()
    flat, slots = dev.step([], 50)  # no creature: nothing to do
    assert flat.size == 0 and slots == []
    with pytest.raises(SimuseError):
        dev.sync_weights([])  # nothing registered: no network to build
    c = creature(20, seed=1)
    bad_col = c.col.copy()
    bad_col[0] = 20
    with pytest.raises(SimuseError):
        dev.register(c.row_ptr, bad_col, c.values, c.nt)
    bad_rp = c.row_ptr.copy()
    bad_rp[0] = 1
    with pytest.raises((SimuseError, ValueError)):
        dev.register(bad_rp, c.col, c.values, c.nt)
    dev.register(c.row_ptr, c.col, c.values, c.nt)
    with pytest.raises(ValueError):
        dev.sync_weights([(np.zeros(c.n_syn + 1, np.float32), True)])
    with pytest.raises(ValueError):
        dev.step([np.zeros(c.n + 1, np.float32)], 10)
    with pytest.raises(SimuseError):
        dev.download_state(3)
    flat, slots = dev.step([None], 3)
    assert flat.size == 3 * c.n and slots[0].n == c.n
    dev.close()
