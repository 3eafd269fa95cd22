"""Small hand-built networks that restate the reference's own test setups.

Each builder returns a NetSpec so the same network runs on the oracle and on
the device.  Path prefix: crates/SiMuse/crates/.
"""
from __future__ import annotations

import numpy as np

from simuse.models import gpea_pop, hh_pop, izh_pop, lif_pop
from simuse.workloads import NetSpec, Source, csr_from_triplets


def delay_net(delays, force=True) -> NetSpec:
    """A[i] -> B[i] with delays[i], weight 500 (clamped to 100), fast LIF
    (inv_tau 0.5, dt 0.5, v_rest -70, v_th -54, tau_ref 2); A forced to fire at
    step 0 with v = 100 (pillars/tests/network_delays.rs:1-99)."""
    n = len(delays)
    v0a = 100.0 if force else -70.0
    a = lif_pop(n, 0.5, 0.5, -70.0, -54.0, -70.0, 2.0, v0a)
    b = lif_pop(n, 0.5, 0.5, -70.0, -54.0, -70.0, 2.0, -70.0)
    pre = np.arange(n, dtype=np.uint32)
    post = pre + n
    rp, col, w, d = csr_from_triplets(2 * n, pre, post, np.full(n, 500.0, np.float32),
                                      np.asarray(delays, np.uint8))
    return NetSpec("delays", [a, b], rp, col, w, d, 0.5, 70)


def two_pop_net(n=8) -> NetSpec:
    """LIF(n) -> Izh(n) w=0.5 d=1 and Izh -> LIF w=0.1 d=0, dt 0.5
    (pillars-cuda/tests/network_parity.rs:19-70)."""
    dt = 0.5
    lif = lif_pop(n, dt, 0.1, -70.0, -54.0, -70.0, 2.0, -70.0)
    izh = izh_pop(n, dt, 0.02, 0.2, -65.0, 8.0, 2.0, 30.0, -65.0, -13.0)
    i = np.arange(n, dtype=np.uint32)
    pre = np.concatenate([i, n + i])
    post = np.concatenate([n + i, i])
    w = np.concatenate([np.full(n, 0.5, np.float32), np.full(n, 0.1, np.float32)])
    d = np.concatenate([np.full(n, 1, np.uint8), np.zeros(n, np.uint8)])
    rp, col, w, d = csr_from_triplets(2 * n, pre, post, w, d)
    return NetSpec("two_pop", [lif, izh], rp, col, w, d, dt, 200)


def empty_csr(n):
    return np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.float32), np.zeros(0, np.uint8)


def random_net(n, k_mean, seed, model="izh", max_delay=0, inh_frac=0.2, noise=5.0, multi_edges=True,
               sources=0, dt=None) -> NetSpec:
    """Random recurrent network with variable out-degree, optional multi-edges
    and delays, Gaussian noise and Poisson sources (test-sized)."""
    rng = np.random.default_rng(seed)
    deg = rng.poisson(k_mean, size=n)
    pre = np.repeat(np.arange(n, dtype=np.uint32), deg)
    post = rng.integers(0, n, size=pre.size).astype(np.uint32)
    if not multi_edges:
        key = pre.astype(np.uint64) * n + post
        _, keep = np.unique(key, return_index=True)
        keep.sort()
        pre, post = pre[keep], post[keep]
    ne = int(n * (1 - inh_frac))
    exc = pre < ne
    if model == "izh":
        w = np.where(exc, rng.random(pre.size) * 0.5, -rng.random(pre.size)).astype(np.float32)
    else:
        w = np.where(exc, rng.random(pre.size) * 20.0, -rng.random(pre.size) * 60.0).astype(np.float32)
    d = rng.integers(0, max_delay + 1, size=pre.size).astype(np.uint8)
    rp, col, w, d = csr_from_triplets(n, pre, post, w, d)
    if model == "izh":
        dt = 1.0 if dt is None else dt
        r = rng.random(n).astype(np.float32)
        isexc = np.arange(n) < ne
        a = np.where(isexc, 0.02, 0.02 + 0.08 * r).astype(np.float32)
        b = np.where(isexc, 0.2, 0.25 - 0.05 * r).astype(np.float32)
        c = np.where(isexc, -65 + 15 * r * r, -65.0).astype(np.float32)
        dd = np.where(isexc, 8 - 6 * r * r, 2.0).astype(np.float32)
        v0 = np.full(n, -65.0, np.float32)
        pop = izh_pop(n, dt, a, b, c, dd, 0.0, 30.0, v0, (b * v0).astype(np.float32))
    elif model == "lif":
        dt = 0.1 if dt is None else dt
        v0 = rng.uniform(-65, -50, n).astype(np.float32)
        pop = lif_pop(n, dt, 0.05, -48.0, -50.0, -65.0, 2.0, v0)  # v_rest above threshold: tonic firing
    elif model == "gpea":
        dt = 0.025 if dt is None else dt
        pop = gpea_pop(n, dt, v0=rng.uniform(-62, -58, n).astype(np.float32))
    else:
        dt = 0.01 if dt is None else dt
        pop = hh_pop(n, dt, v0=rng.uniform(-70, -60, n).astype(np.float32))
    spec = NetSpec(f"random_{model}", [pop], rp, col, w, d, dt, 200)
    if noise:
        spec.noise.append((0, np.full(n, noise, np.float32), seed + 99))
    for s in range(sources):
        m = max(1, n // 50)
        posts = rng.choice(n, size=m, replace=True).astype(np.uint32)
        spec.sources.append(Source(float(rng.uniform(500, 5000)), dt, int(rng.integers(0, 2**32)), posts,
                                   rng.random(m).astype(np.float32) * 5.0,
                                   rng.integers(0, max_delay + 1, m).astype(np.uint8)))
    return spec
