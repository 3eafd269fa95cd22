"""Single-population API (simuse_pop_*, the mirror of pillars_cuda::
CompiledModel, crates/SiMuse/crates/pillars-cuda/src/engine.rs:98-423)
against the oracle (needs an MI355X).

The oracle runs the same population as a network without synapses and the
same current injected every step (what the device keeps resident), so every
step's spike flags and the final state columns must agree bit for bit.
"""
import sys

import numpy as np
import pytest

import pyoracle as po
from nets import empty_csr
from simuse import MODEL_ALPHA_LIF, MODEL_GPEA, MODEL_HH, MODEL_IZHIKEVICH, MODEL_LIF
from simuse.model import CompiledModel
from simuse.models import alpha_lif_pop, gpea_pop, hh_pop, izh_pop, lif_pop, GPEA_PULSE
from simuse.workloads import NetSpec

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle(pop, dt):
    rp, col, w, d = empty_csr(pop.n)
    return po.OracleNetwork(NetSpec("pop", [pop], rp, col, w, d, dt, 0), 32)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _run_against_oracle(pop, dt, cur, steps, every=1, change=None):
    """Persistent path: upload once, step_device per step; spikes of every
    `every`-th step and the final state bit for bit."""
    m = CompiledModel(pop.model, pop.n, pop.stride)
    m.upload(pop.states, pop.params, cur)
    ora = _oracle(pop, dt)
    idx = np.arange(pop.n)
    total = 0
    for t in range(steps):
        if change is not None and t == change[0]:
            cur = change[1]
            m.upload_i_in(cur)
        m.step_device(dt)
        ora.inject_current(0, idx, cur)
        ora.step()
        if t % every == 0:
            s = m.download_spikes()
            assert np.array_equal(s, ora.spikes()[: pop.n].astype(np.int32)), f"step {t}"
            total += int(s.sum())
    st = m.download_states().reshape(pop.ns, pop.stride)
    for c in range(pop.ns):
        assert np.array_equal(_bits(st[c, : pop.n]), _bits(ora.state(0, c))), f"column {c}"
    return m, ora, total


def test_lif_parity_restated():
    # pillars-cuda/tests/lif_parity.rs:26-75: N=512, dt 0.1, I=30 on even / 0 on odd indices, 500 steps
    say("pop lif parity")
    n, dt = 512, 0.1
    pop = lif_pop(n, dt, 0.1, -70.0, -54.0, -70.0, 2.0, -70.0)
    cur = np.where(np.arange(n) % 2 == 0, 30.0, 0.0).astype(np.float32)
    m, ora, total = _run_against_oracle(pop, dt, cur, 500)
    assert total > 0


def test_izhikevich_parity_restated_and_async_steps():
    # pillars-cuda/tests/izhikevich_parity.rs:28-104: N=64, dt 0.5, I=10, 500 steps; then
    # the same 500 steps queued at once (step_graph) end in the same state
    say("pop izhikevich parity")
    n, dt = 64, 0.5
    pop = izh_pop(n, dt, 0.02, 0.2, -65.0, 8.0, 2.0, 30.0, -65.0, -13.0)
    cur = np.full(n, 10.0, np.float32)
    m, ora, total = _run_against_oracle(pop, dt, cur, 500)
    assert total > 0
    a = CompiledModel(pop.model, n, pop.stride)
    a.upload(pop.states, pop.params, cur)
    a.step_graph(dt, 500)
    a.sync()
    assert np.array_equal(_bits(a.download_states()), _bits(m.download_states()))
    assert np.array_equal(a.download_spikes(), m.download_spikes())


def test_hh_gpea_alpha_lif_populations_and_input_change():
    say("pop hh / gpea / alpha-lif")
    rng = np.random.default_rng(3)
    n = 1000
    pop = hh_pop(n, 0.01, v0=rng.uniform(-70.0, -60.0, n).astype(np.float32))
    cur = rng.uniform(0.0, 20.0, n).astype(np.float32)
    _, _, total = _run_against_oracle(pop, 0.01, cur, 600, every=5,
                                      change=(300, rng.uniform(5.0, 30.0, n).astype(np.float32)))
    assert total > 0
    g = gpea_pop(256)
    gcur = np.linspace(0.0, 3.0 * GPEA_PULSE, 256).astype(np.float32)
    _, _, total = _run_against_oracle(g, 0.025, gcur, 2400, every=7)
    assert total > 0
    al = alpha_lif_pop(500)
    acur = rng.uniform(0.0, 200.0, 500).astype(np.float32)
    _, _, total = _run_against_oracle(al, 0.5, acur, 300)
    assert total > 0


def test_compaction_ascending_and_naive_step_equals_persistent():
    say("pop compaction and naive step")
    rng = np.random.default_rng(5)
    n, dt = 70_001, 0.5  # not a multiple of 64
    pop = izh_pop(n, dt, 0.02, 0.2, -65.0, 8.0, 0.0, 30.0, rng.uniform(-70.0, 30.0, n).astype(np.float32), -13.0)
    cur = rng.uniform(0.0, 15.0, n).astype(np.float32)
    m = CompiledModel(pop.model, n, pop.stride)
    m.upload(pop.states, pop.params, cur)
    naive = CompiledModel(pop.model, n, pop.stride)
    st = np.ascontiguousarray(pop.states, np.float32).reshape(-1).copy()
    spk = np.zeros(n, np.int32)
    seen = 0
    for t in range(40):
        m.step_device(dt)
        naive.step(st, pop.params, cur, spk, dt)
        s = m.download_spikes()
        assert np.array_equal(s, spk), t
        c = m.compact_spikes()
        ids = m.download_compact(c)
        assert np.array_equal(ids, np.nonzero(s)[0].astype(np.int32)), t
        seen += c
    assert seen > 0
    assert np.array_equal(_bits(st), _bits(m.download_states()))
    # run(): n steps with the same inputs, states and last spikes back; run(0) touches nothing
    st2 = np.ascontiguousarray(pop.states, np.float32).reshape(-1).copy()
    spk2 = np.zeros(n, np.int32)
    naive.run(st2, pop.params, cur, spk2, dt, 40)
    assert np.array_equal(_bits(st2), _bits(st)) and np.array_equal(spk2, spk)
    before = st2.copy()
    naive.run(st2, pop.params, cur, spk2, dt, 0)
    assert np.array_equal(_bits(st2), _bits(before))
    with pytest.raises(Exception):
        m.download_compact(m.compact_spikes() + 1)
