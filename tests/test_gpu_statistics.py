"""Long-run firing statistics of the device (FP32) against the FP64 oracle
(needs an MI355X).

Over windows long enough for the FP32 and FP64 trajectories to decorrelate,
per-population mean firing rate, mean ISI CV (rates.rs:6-35, isi.rs:30-67) and
the synchrony measure chi = sqrt(Var_t(population mean spike) /
mean_i Var_t(spike_i)) (Golomb 2007, on binary spike trains) must agree within
1%.  The device side uses the device statistics kernel for rates and CVs.
"""
import numpy as np
import pytest

import pyoracle as po
from nets import random_net
from simuse.workloads import Source, build_device, izhikevich_network

pytestmark = pytest.mark.gpu


def _chi(r):
    x = r.astype(np.float64)
    return float(np.sqrt(x.mean(1).var() / max(np.mean(x.var(0)), 1e-30)))


def _drive(spec, w):
    for i in range(spec.n_total):
        spec.sources.append(Source(2000.0, spec.dt, (777 + i * 2654435761) & 0xFFFFFFFF, np.array([i], np.uint32),
                                   np.array([w], np.float32), np.array([0], np.uint8)))
    return spec


def _receptors(spec, g_scale):
    # conductance synapses: excitatory neurons' AMPA, inhibitory ones' GABA_A
    pre = np.repeat(np.arange(spec.n_total), np.diff(spec.row_ptr.astype(np.int64)))
    spec.w = (np.abs(spec.w) * g_scale).astype(np.float32)
    spec.receptors = dict(types=np.where(pre < spec.n_total * 4 // 5, 0, 2).astype(np.uint8), scale=1.0)
    return spec


CASES = {
    "izhikevich": (lambda: izhikevich_network(n=10000, k=100, seed=3), 2000, 1e-3),
    "lif": (lambda: random_net(5000, 50, seed=5, model="lif", max_delay=5, noise=3.0), 3000, 1e-4),
    "hh": (lambda: _drive(random_net(1000, 40, seed=13, model="hh", max_delay=4, noise=40.0), 40.0), 20000, 1e-5),
    "gpea": (lambda: _drive(random_net(1000, 40, seed=13, model="gpea", max_delay=4, noise=60.0), 40.0), 8000,
             2.5e-5),
    "hh_receptors": (lambda: _receptors(_drive(random_net(1000, 40, seed=13, model="hh", max_delay=1, noise=40.0),
                                               40.0), 0.005), 20000, 1e-5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_statistics_match_fp64_oracle_within_1pct(case):
    make, T, dt_s = CASES[case]
    print("statistics", case, flush=True)
    spec = make()
    dev = build_device(spec, graph_steps=16)
    dev.record(T)
    dev.run(T)
    rd = dev.download_raster()
    o = po.OracleNetwork(spec, 64)
    o.record(True)
    o.step(T)
    ro = o.raster().reshape(T, -1)
    off = 0
    for p in spec.pops:
        rates_d, cvs_d = dev.raster_stats(off, p.n, dt_s)
        sl = slice(off, off + p.n)
        rates_o = po.firing_rates(ro[:, sl].astype(np.int32), dt_s)
        cvs_o = po.isi_cv(ro[:, sl].astype(np.int32), dt_s)
        r_d, r_o = float(rates_d.mean()), float(rates_o.mean())
        c_d, c_o = float(cvs_d[cvs_d > 0].mean()), float(cvs_o[cvs_o > 0].mean())
        x_d, x_o = _chi(rd[:, sl]), _chi(ro[:, sl])
        print(f"  pop n={p.n}: rate {r_d:.4f}/{r_o:.4f} cv {c_d:.4f}/{c_o:.4f} chi {x_d:.5f}/{x_o:.5f}", flush=True)
        assert r_o > 1.0
        assert abs(r_d - r_o) <= 0.01 * r_o
        assert abs(c_d - c_o) <= 0.01 * c_o
        assert abs(x_d - x_o) <= 0.01 * x_o
        off += p.n
