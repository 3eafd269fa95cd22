"""Serving runtime: continuous batching with HBM-budget admission (needs an MI355X).

A request's result (steps, spike and event counts; the Kuramoto order
parameter) must be the same when it runs alone and when it shares the device
with other requests that arrive while it runs.
"""
import time

import numpy as np
import pytest

from simuse import MODEL_HH, MODEL_IZHIKEVICH, MODEL_LIF
from simuse.batch import Batch, gen_bytes, gen_desc
from simuse.server import DONE, REJECTED, Server, kuramoto_request, network_request

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def _requests():
    return [
        network_request(gen_desc(MODEL_IZHIKEVICH, 20000, 100, seed=5, dt=1.0, noise_sigma=4.0), 700),
        network_request(gen_desc(MODEL_LIF, 8000, 50, seed=6, dt=0.1, w_exc=20.0, w_inh=60.0, d_min=1, d_max=15,
                                 drive_rate_hz=3000.0, drive_weight=60.0), 300),
        network_request(gen_desc(MODEL_HH, 5000, 40, seed=7, dt=0.01, w_exc=2.0, w_inh=6.0, d_min=1, d_max=4,
                                 noise_sigma=30.0), 500),
        kuramoto_request(20000, modules=16, seed=8, horizon=400),
        network_request(gen_desc(MODEL_IZHIKEVICH, 50000, 100, seed=9, dt=1.0, noise_sigma=5.0,
                                 d_min=0, d_max=3), 1000),
    ]


def _key(res):
    return (res.status, res.steps, res.spikes, res.events, res.order_parameter)


def test_server_result_same_alone_and_in_mix():
    print("server alone vs mix", flush=True)
    reqs = _requests()
    alone = []
    for r in reqs:
        s = Server(8 * GiB, chunk_steps=16)
        rid = s.submit(r)
        s.drain(120.0)
        alone.append(_key(s.result(rid)))
        s.close()
    assert all(a[0] == DONE for a in alone)
    assert all(a[2] > 0 for a, r in zip(alone, reqs) if r.kind == 0)
    # the generated networks' counts equal a plain batch run of the same descriptor
    b = Batch()
    nid = b.add_generated(reqs[0].net)
    b.run(700)
    assert b.counters(nid) == (alone[0][2], alone[0][3])
    b.close()
    # mix: staggered arrivals while earlier requests are on the device
    s = Server(8 * GiB, chunk_steps=8)
    ids = []
    for r in reqs:
        ids.append(s.submit(r))
        time.sleep(0.003)
    s.drain(120.0)
    mixed = [_key(s.result(i)) for i in ids]
    st = s.stats()
    s.close()
    assert mixed == alone
    assert st.peak_live >= 2


def test_server_admission_by_hbm_budget():
    print("server budget", flush=True)
    g = [gen_desc(MODEL_IZHIKEVICH, 30000 + 5000 * i, 100, seed=20 + i, dt=1.0, noise_sigma=5.0) for i in range(6)]
    sizes = [gen_bytes(x) for x in g]
    budget = max(sizes) + min(sizes) + 1   # at most two fit at once
    s = Server(budget, chunk_steps=16)
    ids = [s.submit(network_request(x, 200)) for x in g]
    big = gen_desc(MODEL_IZHIKEVICH, 400000, 100, seed=99, dt=1.0)
    rid_big = s.submit(network_request(big, 10))
    s.drain(120.0)
    st = s.stats()
    res = [s.result(i) for i in ids]
    assert all(r.status == DONE and r.steps == 200 for r in res)
    assert s.result(rid_big).status == REJECTED
    assert st.peak_bytes <= budget and st.peak_live <= 2
    # later requests waited for memory: admitted after earlier ones finished
    assert max(r.t_admit for r in res) >= min(r.t_done for r in res)
    s.close()
