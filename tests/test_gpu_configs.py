"""The benchmark configurations at their stated sizes (needs an MI355X).

* Brunel LIF: 10k neurons, p = 0.01, delays 1..15, 2 kHz Poisson per neuron,
  dt 0.1 ms, the whole 10,000-step horizon bit for bit against the oracle.
* Hodgkin-Huxley: 200k neurons, Pareto out-degree (~40M synapses), lognormal
  conductance synapses and pair STDP, bit for bit against the oracle's FP32
  twin, and the FP32-vs-FP64 tolerance from a common start.
* Hodgkin-Huxley over the stated 50,000-step horizon (a 10k-neuron network of
  the same construction, so that the oracle's FP64 twin finishes in seconds):
  firing rate, ISI CV, synchrony and the mean plastic weight within 1% of
  FP64.
* Batching: the full 512-network mix; a sample of networks equal alone and in
  the mix, generated networks equal to the oracle.
* Serving: the 64-request Poisson stream at 200 requests/s under an HBM budget
  small enough that requests wait for memory; a sample of results equal to the
  same requests served alone.
"""
import os
import tempfile
import time

import numpy as np
import pytest

import pyoracle as po
from simuse import MODEL_IZHIKEVICH, MODEL_LIF
from simuse.batch import Batch, gen_bytes
from simuse.server import DONE, REQ_KURAMOTO, Server
from simuse.workloads import (batching_mix, brunel_lif_network, build_device, hh_network, serving_mix,
                              spec_from_generated)

pytestmark = pytest.mark.gpu

THREADS = 16


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same_state(dev, ora, spec):
    for p, pp in enumerate(spec.pops):
        d = dev.download_states(p)
        for c in range(d.shape[0]):
            assert np.array_equal(_bits(d[c, : pp.n]), _bits(ora.state(p, c))), (p, c)


def test_brunel_lif_stated_size_whole_horizon_bitexact():
    print("config brunel 10k p=0.01 2kHz 10000 steps", flush=True)
    spec = brunel_lif_network(n=10_000, p=0.01, seed=1, steps=10_000, rate_hz=2000.0)
    assert spec.n_total == 10_000 and 0.9e6 < spec.n_syn < 1.1e6 and len(spec.sources) == 10_000
    dev = build_device(spec, graph_steps=16)
    ora = po.OracleNetwork(spec, 32)
    spikes = 0
    for t in range(2000):  # spike ids every step
        dev.step()
        ora.step()
        ids = dev.download_spike_ids()
        assert np.array_equal(ids, ora.active()), t
        spikes += ids.size
    for _ in range(8):  # then 1000-step chunks through the step graphs
        dev.run(1000)
        ora.step(1000)
        assert dev.counters() == ora.counters()
        _same_state(dev, ora, spec)
    assert np.array_equal(dev.download_ring().view(np.uint32), ora.ring().astype(np.float32).view(np.uint32))
    s, e = ora.counters()
    print(f"  {s} spikes ({spikes} in the first 2000 steps), {e} events over 10000 steps", flush=True)
    assert e > 10_000 * 1500  # the Poisson drive delivers ~2000 events per step
    dev.close()


def test_hh_200k_conductances_stdp_stated_size():
    print("config hh 200k receptors stdp", flush=True)
    spec = hh_network(n=200_000, seed=3, steps=600, noise_sd=30.0, receptors=True)
    assert spec.n_total == 200_000 and 3.5e7 < spec.n_syn < 4.5e7
    dev = build_device(spec, graph_steps=16)
    ora = po.OracleNetwork(spec, 32, threads=THREADS)
    w0 = spec.w.copy()
    for t in range(400):
        dev.step()
        ora.step()
        assert np.array_equal(dev.download_spike_ids(), ora.active()), t
    _same_state(dev, ora, spec)
    assert dev.counters() == ora.counters()
    w = dev.download_weights()
    assert np.array_equal(_bits(w), _bits(ora.weights()))
    assert np.array_equal(_bits(dev.download_conductances()), _bits(ora.conductances()))
    s, e = ora.counters()
    changed = int(np.count_nonzero(w != w0))
    print(f"  {s} spikes, {e} events, {changed} weights changed by STDP over 400 steps", flush=True)
    assert s > 0 and changed > 0
    ora.close()
    # FP32 vs FP64 from a common start (the device snapshot), 100 steps = 1 ms
    snap = os.path.join(tempfile.mkdtemp(), "hh200k.taas")
    dev.save_state(snap)
    o64 = po.OracleNetwork(spec, 64, threads=THREADS)
    o64.set_state_from_snapshot(snap)
    os.remove(snap)
    dev.run(100)
    o64.step(100)
    dv = np.abs(dev.download_states(0)[0, :200_000].astype(np.float64) - o64.state(0, 0))
    dg = np.abs(dev.download_conductances().astype(np.float64) - o64.conductances())
    print(f"  FP64 window: |dv| mean {dv.mean():.2e} max {dv.max():.2e} mV, |dg| max {dg.max():.2e}", flush=True)
    assert dv.mean() < 2e-4 and dv.max() < 5e-3 and dg.max() < 1e-3
    o64.close()
    dev.close()


class _SpikeStats:
    """Rate, ISI CV (P/src/kernels/stats/rates.rs:6-35, isi.rs:30-67: n-1
    variance, at least two intervals) and the synchrony chi of a raster fed
    in consecutive chunks, without holding the whole raster."""

    def __init__(self, n):
        self.n = n
        self.count = np.zeros(n, np.int64)
        self.last = np.full(n, -1, np.int64)
        self.s1 = np.zeros(n)
        self.s2 = np.zeros(n)
        self.k = np.zeros(n, np.int64)
        self.pop = []
        self.steps = 0

    def add(self, r):
        r = np.asarray(r)[:, : self.n]
        t0 = self.steps
        self.steps += r.shape[0]
        self.pop.append(r.sum(1, dtype=np.int64))
        self.count += r.sum(0, dtype=np.int64)
        t, i = np.nonzero(r)
        order = np.lexsort((t, i))  # by neuron, then time
        t, i = t[order].astype(np.int64) + t0, i[order]
        if t.size == 0:
            return
        first = np.ones(t.size, bool)
        first[1:] = i[1:] != i[:-1]
        prev = np.empty_like(t)
        prev[1:] = t[:-1]
        prev[first] = self.last[i[first]]
        ok = prev >= 0
        isi = (t - prev)[ok].astype(np.float64)
        np.add.at(self.s1, i[ok], isi)
        np.add.at(self.s2, i[ok], isi * isi)
        np.add.at(self.k, i[ok], 1)
        lastm = np.ones(t.size, bool)
        lastm[:-1] = i[:-1] != i[1:]
        self.last[i[lastm]] = t[lastm]

    def result(self, dt_s):
        T = self.steps
        rate = float(self.count.mean() / (T * dt_s))
        m = self.k >= 2
        mean = self.s1[m] / self.k[m]
        var = (self.s2[m] - self.k[m] * mean * mean) / (self.k[m] - 1)
        cv = float(np.mean(np.sqrt(np.maximum(var, 0.0)) / mean))
        p = self.count / T
        pop = np.concatenate(self.pop) / self.n
        chi = float(np.sqrt(pop.var() / max(np.mean(p * (1 - p)), 1e-30)))
        return rate, cv, chi


def test_hh_conductances_stdp_full_horizon_statistics():
    # The precision network's construction (Pareto out-degree, lognormal AMPA/GABA_A
    # conductances, pair STDP with FP64 weight masters) over its whole
    # 50,000-step horizon: the FP32 device against the oracle's FP64 twin
    # from the same start.  The trajectories decorrelate within a few
    # thousand steps (spike-time flips), so the criterion is statistical.
    print("config hh full horizon statistics", flush=True)
    T, chunk, n = 50_000, 1000, 10_000
    spec = hh_network(n=n, seed=3, steps=T, noise_sd=30.0, receptors=True)
    dev = build_device(spec, graph_steps=16)
    o64 = po.OracleNetwork(spec, 64, threads=THREADS)
    sd, so = _SpikeStats(n), _SpikeStats(n)
    t0 = time.perf_counter()
    for c in range(T // chunk):
        dev.record(chunk)
        dev.run(chunk)
        sd.add(dev.download_raster())
        o64.record(True)
        o64.step(chunk)
        so.add(o64.raster().reshape(chunk, -1))
        if c % 10 == 9:
            print(f"  {c + 1} chunks, {time.perf_counter() - t0:.1f} s", flush=True)
    dt_s = spec.dt * 1e-3
    (rd, cd, xd), (ro, co, xo) = sd.result(dt_s), so.result(dt_s)
    plastic = spec.stdp["plastic"].astype(bool)
    wd = float(dev.download_weights().astype(np.float64)[plastic].mean())
    wo = float(o64.weights()[plastic].mean())
    w0 = float(spec.w.astype(np.float64)[plastic].mean())
    print(f"  rate {rd:.4f}/{ro:.4f} Hz, cv {cd:.4f}/{co:.4f}, chi {xd:.5f}/{xo:.5f}, "
          f"mean plastic weight {wd:.6f}/{wo:.6f} (start {w0:.6f})", flush=True)
    assert ro > 1.0 and sd.steps == T
    assert abs(rd - ro) <= 0.01 * ro
    assert abs(cd - co) <= 0.01 * co
    assert abs(xd - xo) <= 0.01 * xo
    assert abs(wo - w0) > 1e-4  # STDP moved the weights
    assert abs(wd - wo) <= 0.01 * abs(wo)
    o64.close()
    dev.close()


def test_batching_full_mix_alone_equals_batched_and_oracle():
    print("config batching 512 networks", flush=True)
    descs = batching_mix(512, seed=1)
    assert len(descs) == 512 and 1.5e7 < sum(d.n for d in descs) < 2.5e7
    T = 200
    b = Batch()
    ids = [b.add_generated(d) for d in descs]
    b.set_graph_steps(16)
    b.autotune(16)  # schedule only: no effect on results
    b.run(T)
    rng = np.random.default_rng(7)
    sample = sorted(int(i) for i in rng.choice(512, size=8, replace=False))
    mixed = {}
    for i in sample:
        mixed[i] = (b.counters(ids[i]), b.download_spike_ids(ids[i]),
                    [b.download_state(ids[i], p) for p in range(b.info(ids[i])[2])])
    b.close()
    for i in sample:
        a = Batch()
        nid = a.add_generated(descs[i])
        a.set_graph_steps(16)
        a.run(T)
        cnt, spk, st = mixed[i]
        assert a.counters(nid) == cnt, i
        assert np.array_equal(a.download_spike_ids(nid), spk), i
        for p, s in enumerate(st):
            assert np.array_equal(_bits(a.download_state(nid, p)), _bits(s)), (i, p)
        a.close()
    # generated networks of the mix against the oracle: the smallest LIF and
    # Izhikevich networks and one of about 50k neurons
    lif = min((d.n, k) for k, d in enumerate(descs) if d.model == MODEL_LIF)[1]
    izh = min((d.n, k) for k, d in enumerate(descs) if d.model == MODEL_IZHIKEVICH)[1]
    mid = min((abs(d.n - 50_000), k) for k, d in enumerate(descs))[1]
    for k in (lif, izh, mid):
        d = descs[k]
        a = Batch()
        nid = a.add_generated(d)
        spec = spec_from_generated(a, nid, d, T)
        ora = po.OracleNetwork(spec, 32, threads=THREADS)
        a.set_graph_steps(16)
        for t in range(T):
            a.run(1)
            ora.step()
            assert np.array_equal(a.download_spike_ids(nid), ora.active()), (k, t)
        for p, pp in enumerate(spec.pops):
            st = a.download_state(nid, p)
            for c in range(st.shape[0]):
                assert np.array_equal(_bits(st[c, : pp.n]), _bits(ora.state(p, c))), (k, p, c)
        assert a.counters(nid) == ora.counters()
        print(f"  network {k} (model {d.model}, n={d.n}) = oracle, {ora.counters()[0]} spikes", flush=True)
        ora.close()
        a.close()


def _key(res):
    return (res.status, res.steps, res.spikes, res.events, res.order_parameter)


def test_serving_stream_under_tight_budget_same_as_alone():
    print("config serving 64 requests at 200/s, tight budget", flush=True)
    reqs = serving_mix(64, rate_hz=200.0, seed=3)
    sizes = []
    for _, r in reqs:
        sizes.append(gen_bytes(r.net) if r.kind != REQ_KURAMOTO else None)
    budget = 6 << 30
    srv = Server(budget, chunk_steps=16)
    ids = []
    t0 = time.perf_counter()
    for t, r in reqs:
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        ids.append(srv.submit(r))
    srv.drain(110.0)
    res = [srv.result(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert all(x.status == DONE for x in res), [x.status for x in res]
    assert all(x.steps == r.horizon for x, (_, r) in zip(res, reqs))
    wait = np.array([x.t_admit - x.t_submit for x in res])
    print(f"  peak {st.peak_bytes / 2**30:.2f} GiB of {budget / 2**30:.0f}, peak live {st.peak_live}, "
          f"admission wait max {wait.max():.3f} s, mean {wait.mean():.3f} s", flush=True)
    assert st.peak_bytes <= budget
    # the budget made requests wait: some were admitted only after others finished
    assert max(x.t_admit for x in res) > min(x.t_done for x in res)
    assert wait.max() > 0.05
    # a sample (the four network requests and the Kuramoto request with the
    # least work) served alone gives the same results
    nets = sorted((r.net.n * r.horizon, k) for k, (_, r) in enumerate(reqs) if r.kind != REQ_KURAMOTO)[:4]
    kur = sorted((r.kura.n * r.horizon, k) for k, (_, r) in enumerate(reqs) if r.kind == REQ_KURAMOTO)[:1]
    for _, k in nets + kur:
        s = Server(budget, chunk_steps=16)
        rid = s.submit(reqs[k][1])
        s.drain(60.0)
        alone = _key(s.result(rid))
        s.close()
        assert alone == _key(res[k]), k
