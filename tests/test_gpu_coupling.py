"""Electrical coupling on the MI355X against the oracle (needs an MI355X).

Gap junctions, graded and sigmoid graded synapses on random two-population
networks, in the binned two-kernel form (every delay <= 1) and in the
chained form (delays up to 5), with Poisson sources and Gaussian noise: spike
ids every step, then state, ring and the sigmoid activations through a
checkpoint, bit for bit.  Then the boundary's refusals.
"""
import os
import tempfile

import numpy as np
import pytest

import pyoracle as po
from nets import random_net
from simuse import SimuseError
from simuse._lib import E_INVALID, E_STATE, E_UNSUPPORTED, FORM_ONE_KERNEL
from simuse.models import make_pop
from simuse.network import N_PARAMS, N_STATE
from simuse.workloads import build_device

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _split(spec, k):
    """The single population as two: neurons [0, k) and [k, n)."""
    p = spec.pops[0]
    n = p.n
    cols = lambda a, lo, hi: [a[c, lo:hi] for c in range(a.shape[0])]  # noqa: E731
    spec.pops = [make_pop(p.model, k, cols(p.params, 0, k), cols(p.states, 0, k)),
                 make_pop(p.model, n - k, cols(p.params, k, n), cols(p.states, k, n))]
    assert len(spec.pops[0].params) == N_PARAMS[p.model] and len(spec.pops[0].states) == N_STATE[p.model]
    spec.noise = [(0, sig[:k], seed) for _, sig, seed in spec.noise]
    return spec


def _coupled(max_delay, seed):
    spec = _split(random_net(1800, 20, seed=seed, model="izh", max_delay=max_delay, noise=6.0, sources=3), 1300)
    n = spec.n_total
    rng = np.random.default_rng(seed)
    gap = (rng.integers(0, n, 2500), rng.integers(0, n, 2500), rng.uniform(0.0, 0.08, 2500).astype(np.float32))
    graded = (rng.integers(0, n, 1200), rng.integers(0, n, 1200), rng.uniform(0.0, 0.1, 1200).astype(np.float32))
    sig = [((-55.0, 4.0, 0.4, 0.0, 0.05, 1.0),
            (rng.integers(0, n, 700), rng.integers(0, n, 700), rng.uniform(0.2, 1.0, 700).astype(np.float32))),
           ((-60.0, -3.0, 0.8, -80.0, 0.02, 1.0),
            (rng.integers(0, n, 500), rng.integers(0, n, 500), rng.uniform(0.2, 1.0, 500).astype(np.float32)))]
    spec.coupling = dict(gap=gap, graded=graded, sigmoid=sig)
    return spec


@pytest.mark.parametrize("max_delay", [1, 5])
def test_coupled_network_bitexact_with_checkpoint(max_delay):
    print(f"coupling: izh 1800 (two populations), gap/graded/sigmoid, max delay {max_delay}", flush=True)
    spec = _coupled(max_delay, seed=11 + max_delay)
    dev = build_device(spec, graph_steps=8)
    assert dev.step_form() != FORM_ONE_KERNEL
    ora = po.OracleNetwork(spec, 32)
    spikes = 0
    for t in range(150):
        dev.step()
        ora.step()
        ids = dev.download_spike_ids()
        assert np.array_equal(ids, ora.active()), t
        spikes += ids.size
    snap = os.path.join(tempfile.mkdtemp(), "coupled.taas")
    dev.save_state(snap)
    dev.run(150)
    ora.step(150)
    for p, pp in enumerate(spec.pops):
        d = dev.download_states(p)
        for c in range(d.shape[0]):
            assert np.array_equal(_bits(d[c, : pp.n]), _bits(ora.state(p, c))), (p, c)
    assert dev.counters() == ora.counters()
    assert np.array_equal(dev.download_ring().view(np.uint32), ora.ring().astype(np.float32).view(np.uint32))
    # resume: a network built the same way continues from the snapshot bit for bit
    again = build_device(spec, graph_steps=8)
    again.load_state(snap)
    os.remove(snap)
    again.run(150)
    for p in range(len(spec.pops)):
        assert np.array_equal(_bits(again.download_states(p)), _bits(dev.download_states(p))), p
    print(f"  {spikes} spikes in 150 steps, {ora.counters()} after 300", flush=True)
    assert spikes > 1000
    again.close()
    dev.close()


def test_coupling_refusals():
    spec = random_net(300, 10, seed=2, model="lif", max_delay=1, noise=0.0)
    net = build_device(spec)
    with pytest.raises(SimuseError) as e:
        net.connect_gap_junction(0, 0, 0, 1, float("nan"))
    assert e.value.code == E_INVALID
    with pytest.raises(SimuseError) as e:
        net.connect_graded_synapse(0, 0, 0, 300, 0.1)
    assert e.value.code == E_INVALID
    with pytest.raises(SimuseError) as e:
        net.register_sigmoid_graded(-50.0, 0.0, 0.5, 0.0, 1.0, 0.1)  # delta 0
    assert e.value.code == E_INVALID
    import ctypes as C
    from simuse._lib import lib
    pid = C.c_int32()
    assert lib().simuse_net_register_kinetics(net._h, 0, None, 0, 0.1, C.byref(pid)) == E_UNSUPPORTED
    net.connect_gap_junction(0, 0, 0, 1, 0.1)
    net.step()
    with pytest.raises(SimuseError) as e:
        net.connect_gap_junction(0, 2, 0, 3, 0.1)
    assert e.value.code == E_STATE
    net.close()
    # coupling with receptor synapses, or with the fused Poisson drive: refused at the first step
    rec = build_device(spec)
    rec.set_receptors(np.zeros(spec.n_syn, np.uint8), 1.0, None)
    rec.connect_gap_junction(0, 0, 0, 1, 0.1)
    with pytest.raises(SimuseError) as e:
        rec.step()
    assert e.value.code == E_UNSUPPORTED
    rec.close()
    drv = build_device(spec)
    drv.set_poisson_drive(0, 1000.0, 0.1, 2.0, 7)
    drv.connect_graded_synapse(0, 0, 0, 1, 0.1)
    with pytest.raises(SimuseError) as e:
        drv.step()
    assert e.value.code == E_UNSUPPORTED
    drv.close()
