"""Kuramoto coupling on the device against the FP64 oracle (needs an MI355X).

Tolerances: FP32 sin/cos and derivatives with an FP64 phase give phase
errors of order 1e-6 rad per unit time; the tests allow 1e-4 rad.
"""
import numpy as np
import pytest

import pyoracle as po
from simuse.kuramoto import KuramotoGraph
from simuse.workloads import modular_small_world

pytestmark = pytest.mark.gpu


def wrapped_err(a, b):
    return float(np.max(np.abs(np.angle(np.exp(1j * (a - b))))))


@pytest.mark.parametrize("method,lanes", [("rk4", None), ("rk4", 1), ("rk4", 16), ("euler", None)])
def test_kuramoto_matches_fp64_oracle(method, lanes):
    print("kuramoto", method, lanes, flush=True)
    rp, col, w, om, th = modular_small_world(20_000, seed=3)
    dt, steps = 0.001, 500
    dev = KuramotoGraph(rp, col, w, om, th, dt, method)
    if lanes:
        dev.set_group(lanes)
    ora = po.OracleKuramoto(rp, col, w, om, th, dt, kind=64)
    dev.run(steps)
    (ora.rk4 if method == "rk4" else ora.euler)(steps)
    err = wrapped_err(dev.phases(), ora.phases())
    print("max phase error", err, flush=True)
    assert err < 1e-4
    assert abs(dev.order_parameter() - ora.order()) < 1e-6
    assert dev.step_counter() == steps


def test_kuramoto_identical_frequencies_synchronise():
    n = 4096
    rp, col, w, _, th = modular_small_world(n, modules=16, k=20, beta=0.3, inter=0.1, coupling=4.0, seed=1)
    dev = KuramotoGraph(rp, col, w, np.full(n, 0.5, np.float32), th, 0.01, "rk4")
    r0 = dev.order_parameter()
    dev.run(3000)
    assert dev.order_parameter() > 0.99 > r0


def test_kuramoto_uncoupled_linear_phase():
    n = 1000
    rng = np.random.default_rng(0)
    om = rng.normal(size=n).astype(np.float32)
    th = rng.uniform(-3, 3, n).astype(np.float32)
    rp = np.zeros(n + 1, np.uint32)
    dev = KuramotoGraph(rp, np.zeros(0, np.uint32), np.zeros(0, np.float32), om, th, 0.01, "rk4")
    dev.run(1000)
    t = 1000 * float(np.float32(0.01))
    want = th.astype(np.float64) + om.astype(np.float64) * t
    assert wrapped_err(dev.phases(), want) < 1e-9


def test_phase_coherence_weighted():
    # phase_coherence (pillars/src/kernels/measure/mod.rs:76-106) against the FP64 formula on the
    # device's own phases; unit amplitudes reduce to the order parameter
    n = 5000
    rp, col, w, om, th = modular_small_world(n, modules=8, seed=4)
    dev = KuramotoGraph(rp, col, w, om, th, 0.01, "rk4")
    dev.run(50)
    p = dev.phases()
    a = np.random.default_rng(1).uniform(0.0, 2.0, n).astype(np.float32)
    want = abs(np.sum(a.astype(np.float64) * np.exp(1j * p))) / np.sum(a.astype(np.float64))
    assert abs(dev.phase_coherence(a) - want) < 1e-9
    assert abs(dev.phase_coherence(np.ones(n, np.float32)) - dev.order_parameter()) < 1e-12
    assert dev.phase_coherence(np.zeros(n, np.float32)) == 0.0


@pytest.mark.parametrize("method", ["rk4", "euler"])
def test_kuramoto_row_weight_path_bitexact(method):
    # The modular small world's rows each carry one weight (K / k_i), so the
    # device reads it once per row.  The same graph with one more oscillator
    # that no one listens to, whose two edges have different weights, makes
    # every row read its weights per edge (and its extent from row_ptr):
    # the first n phases must stay bit for bit the same.
    print("kuramoto row-weight path", method, flush=True)
    n = 8192
    rp, col, w, om, th = modular_small_world(n, modules=8, seed=6)
    a = KuramotoGraph(rp, col, w, om, th, 0.01, method)
    rp2 = np.concatenate([rp, [rp[-1] + 2]]).astype(np.uint32)
    col2 = np.concatenate([col, [0, 1]]).astype(np.uint32)
    w2 = np.concatenate([w, [0.25, 0.5]]).astype(np.float32)
    om2 = np.concatenate([om, [0.3]]).astype(np.float32)
    th2 = np.concatenate([th, [0.1]]).astype(np.float32)
    b = KuramotoGraph(rp2, col2, w2, om2, th2, 0.01, method)
    a.set_group(8)  # the same lanes per row, so the same summation order
    b.set_group(8)
    a.run(300)
    b.run(300)
    pa, pb = a.phases(), b.phases()[:n]
    assert np.array_equal(pa.view(np.uint64), pb.view(np.uint64))


@pytest.mark.parametrize("method,lanes", [("rk4", None), ("rk4", 1), ("rk4", 2), ("euler", 4)])
def test_kuramoto_ragged_rows_match_fp64_oracle(method, lanes):
    """Rows of degree 0 to 70 with a weight per edge: empty rows, rows shorter
    than one batch of a lane's edges and rows of several batches."""
    print("kuramoto ragged", method, lanes, flush=True)
    rng = np.random.default_rng(11)
    n = 3000
    deg = rng.integers(0, 71, size=n)
    deg[:50] = 0
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum(deg)
    col = rng.integers(0, n, size=int(rp[-1])).astype(np.uint32)
    w = (rng.random(col.size) * 0.2 - 0.05).astype(np.float32)
    om = rng.normal(0.0, 1.0, n).astype(np.float32)
    th = rng.uniform(-np.pi, np.pi, n).astype(np.float32)
    dt, steps = 0.001, 400
    dev = KuramotoGraph(rp, col, w, om, th, dt, method)
    if lanes:
        dev.set_group(lanes)
    ora = po.OracleKuramoto(rp, col, w, om, th, dt, kind=64)
    dev.run(steps)
    (ora.rk4 if method == "rk4" else ora.euler)(steps)
    err = wrapped_err(dev.phases(), ora.phases())
    print("max phase error", err, flush=True)
    assert err < 1e-4
    assert dev.step_counter() == steps
