"""Seeded synthetic networks for the benchmark configurations and tests.

A :class:`NetSpec` holds everything needed to build the same network on the
device and in a CPU checker: populations, the CSR, Poisson sources and
Gaussian current noise.  All randomness comes from numpy's seeded PCG64, so a
spec is reproducible from its seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from .models import alpha_lif_pop, hh_pop, izh_pop, lif_pop
from .network import CompiledNetwork, NetworkPop


@dataclass
class Source:
    rate_hz: float
    dt_ms: float
    seed: int
    posts: np.ndarray
    weights: np.ndarray
    delays: np.ndarray


@dataclass
class NetSpec:
    name: str
    pops: List[NetworkPop]
    row_ptr: np.ndarray
    col: np.ndarray
    w: np.ndarray
    delay: np.ndarray
    dt: float
    steps: int
    sources: List[Source] = field(default_factory=list)
    noise: List[Tuple[int, np.ndarray, int]] = field(default_factory=list)
    stdp: Optional[dict] = None
    # Tsodyks-Markram short-term plasticity: {profiles: [(u0, tau_rec, tau_fac)], dt,
    # syn_types: uint8[n_syn] (0 = static), src_types: uint8[source edges] or absent}
    stp: Optional[dict] = None
    # conductance-based receptor synapses: {types: uint8[n_syn] (0..6, 255 current-based),
    # scale: float, ext: float32[n_total] or absent}
    receptors: Optional[dict] = None
    # electrical coupling, global indices in call order: {gap: (a, b, g), graded: (pre, post, g),
    # sigmoid: [((vth, delta, k, erev, conductance, dt), (pre, post, weight)), ...]}
    coupling: Optional[dict] = None

    @property
    def n_total(self) -> int:
        return int(sum(p.n for p in self.pops))

    @property
    def n_syn(self) -> int:
        return int(self.col.size)


def csr_from_triplets(n: int, pre, post, w, d):
    """CSR from (pre, post, w, delay) with a *stable* sort by pre.

    The reference sorts with ``sort_unstable_by_key`` (P/src/api/net/graph.rs:147,
    crates/SiMuse/crates/pillars/ prefix), which leaves the order of a row's
    edges unspecified; a stable sort keeps insertion order, and both the device
    and the checker then consume the same CSR arrays.
    """
    pre = np.asarray(pre, dtype=np.uint32)
    order = np.argsort(pre, kind="stable")
    counts = np.bincount(pre, minlength=n).astype(np.uint64)
    rp = np.zeros(n + 1, dtype=np.uint64)
    np.cumsum(counts, out=rp[1:])
    return (rp.astype(np.uint32), np.asarray(post, dtype=np.uint32)[order],
            np.asarray(w, dtype=np.float32)[order], np.asarray(d, dtype=np.uint8)[order])


def _fixed_outdegree_targets(rng: np.random.Generator, n: int, k: int) -> np.ndarray:
    """k distinct targets per neuron, uniform, no autapses; rows sorted."""
    tgt = rng.integers(0, n - 1, size=(n, k), dtype=np.int64)
    pre = np.arange(n, dtype=np.int64)[:, None]
    tgt += tgt >= pre
    tgt.sort(axis=1)
    for _ in range(64):
        dup = np.zeros(tgt.shape, dtype=bool)
        dup[:, 1:] = tgt[:, 1:] == tgt[:, :-1]
        rows = np.nonzero(dup.any(axis=1))[0]
        if rows.size == 0:
            break
        sub = tgt[rows]
        dsub = dup[rows]
        new = rng.integers(0, n - 1, size=int(dsub.sum()), dtype=np.int64)
        rr = np.broadcast_to(rows[:, None], dsub.shape)[dsub]
        new += new >= rr
        sub[dsub] = new
        sub.sort(axis=1)
        tgt[rows] = sub
    return tgt.astype(np.uint32)


def izhikevich_network(n: int = 1_000_000, k: int = 100, seed: int = 1, steps: int = 1000,
                       noise: bool = True) -> NetSpec:
    """Izhikevich (2003) network: 80% regular spiking, 20% fast spiking.

    Two populations (RS then FS).  RS: a=0.02, b=0.2, c=-65+15r^2,
    d=8-6r^2; FS: a=0.02+0.08r, b=0.25-0.05r, c=-65, d=2; r ~ U(0,1).  k
    random distinct targets per neuron, weights exc U(0,0.5), inh U(-1,0),
    delay 0 (next step).  dt = 1 ms integrated as two 0.5 ms half steps by the
    model; no refractory period; threshold 30; v0=-65, u0=b*v0.  Thalamic
    input: Gaussian current, sd 5 (RS) and 2 (FS) per step.
    """
    rng = np.random.default_rng(seed)
    ne = n * 4 // 5
    ni = n - ne
    re = rng.random(ne, dtype=np.float32)
    ri = rng.random(ni, dtype=np.float32)
    v0e = np.full(ne, -65.0, np.float32)
    v0i = np.full(ni, -65.0, np.float32)
    ce = (-65.0 + 15.0 * re * re).astype(np.float32)
    de = (8.0 - 6.0 * re * re).astype(np.float32)
    ai = (0.02 + 0.08 * ri).astype(np.float32)
    bi = (0.25 - 0.05 * ri).astype(np.float32)
    rs = izh_pop(ne, 1.0, 0.02, 0.2, ce, de, 0.0, 30.0, v0e, np.float32(0.2) * v0e)
    fs = izh_pop(ni, 1.0, ai, bi, -65.0, 2.0, 0.0, 30.0, v0i, (bi * v0i).astype(np.float32))
    tgt = _fixed_outdegree_targets(rng, n, k)
    w = np.empty((n, k), dtype=np.float32)
    w[:ne] = rng.random((ne, k), dtype=np.float32) * np.float32(0.5)
    w[ne:] = -rng.random((ni, k), dtype=np.float32)
    rp = (np.arange(n + 1, dtype=np.uint64) * k).astype(np.uint32)
    spec = NetSpec("izhikevich", [rs, fs], rp, tgt.reshape(-1), w.reshape(-1),
                   np.zeros(n * k, dtype=np.uint8), 1.0, steps)
    if noise:
        spec.noise.append((0, np.full(ne, 5.0, np.float32), seed * 7919 + 17))
        spec.noise.append((1, np.full(ni, 2.0, np.float32), seed * 7919 + 18))
    return spec


def brunel_lif_network(n: int = 10_000, p: float = 0.01, seed: int = 1, steps: int = 10_000,
                       rate_hz: float = 2000.0) -> NetSpec:
    """Brunel-style current-based LIF network (80% exc / 20% inh).

    Erdős–Rényi connectivity with probability p, no autapses; J_E=+0.1 mV and
    J_I=-0.5 mV expressed as the current the forward-Euler LIF integrates for
    one step (i = J*tau_m/dt: +20 and -100 at tau_m=20 ms, dt=0.1 ms); delays
    uniform 1..15 steps; tau_m=20 ms, V_rest=V_reset=-65, V_th=-50,
    refractory 2 ms; one independent Poisson source per neuron at rate_hz
    with weight J_E; v0 ~ U(-65,-50).
    """
    rng = np.random.default_rng(seed)
    dt = 0.1
    tau_m = 20.0
    ne = n * 4 // 5
    # Bernoulli(p) per ordered pair, drawn row by row
    pres, posts = [], []
    for i in range(n):
        m = rng.random(n) < p
        m[i] = False
        j = np.nonzero(m)[0]
        pres.append(np.full(j.size, i, np.uint32))
        posts.append(j.astype(np.uint32))
    pre = np.concatenate(pres)
    post = np.concatenate(posts)
    je = np.float32(0.1 * tau_m / dt)
    ji = np.float32(-0.5 * tau_m / dt)
    w = np.where(pre < ne, je, ji).astype(np.float32)
    d = rng.integers(1, 16, size=pre.size).astype(np.uint8)
    rp, col, w, d = csr_from_triplets(n, pre, post, w, d)
    v0 = rng.uniform(-65.0, -50.0, size=n).astype(np.float32)
    pop = lif_pop(n, dt, 1.0 / tau_m, -65.0, -50.0, -65.0, 2.0, v0)
    spec = NetSpec("brunel_lif", [pop], rp, col, w, d, dt, steps)
    seeds = (np.arange(n, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(seed)) & np.uint64(0xFFFFFFFF)
    for i in range(n):
        spec.sources.append(Source(rate_hz, dt, int(seeds[i]), np.array([i], np.uint32),
                                   np.array([je], np.float32), np.array([0], np.uint8)))
    return spec


def hh_network(n: int = 200_000, seed: int = 1, steps: int = 50_000, mean_deg: float = 200.0,
               w_scale: float = 1.0, noise_sd: float = 3.0, stdp: bool = True, receptors: bool = False,
               g_unit: float = 0.003) -> NetSpec:
    """Hodgkin–Huxley network with heavy-tailed out-degree and pair STDP.

    Out-degree ceil(x_m * U^(-1/2)) (Pareto alpha=2, x_m = mean/2, mean 200),
    capped at n-1; excitatory weights lognormal(mu=-1, sigma=1) scaled by
    w_scale, 20% inhibitory neurons with weight -lognormal; dt = 0.01 ms;
    Gaussian current noise keeps the network active.  STDP (A+=0.01,
    A-=0.012, tau=20 ms, clip [0,1]) on excitatory synapses.
    Without `receptors` the weights are current pulses through the ring;
    with it they are conductances: excitatory synapses AMPA, inhibitory ones
    GABA_A (the lognormal magnitude), each spike jumping g by weight * g_unit
    (mS/cm^2; 0.003 keeps the network asynchronous at about 10 Hz).
    """
    rng = np.random.default_rng(seed)
    dt = 0.01
    xm = mean_deg / 2.0
    u = rng.random(n)
    deg = np.minimum(np.ceil(xm * u ** -0.5), n - 1).astype(np.int64)
    rp = np.zeros(n + 1, dtype=np.uint64)
    np.cumsum(deg, out=rp[1:])
    ns = int(rp[-1])
    pre = np.repeat(np.arange(n, dtype=np.int64), deg)
    post = rng.integers(0, n - 1, size=ns, dtype=np.int64)
    post += post >= pre
    ne = n * 4 // 5
    w = rng.lognormal(-1.0, 1.0, size=ns).astype(np.float32) * np.float32(w_scale)
    exc = pre < ne
    w = np.where(exc, np.minimum(w, 1.0), -w).astype(np.float32)
    pop = hh_pop(n, dt, tau_ref=2.0, v0=rng.uniform(-70.0, -60.0, size=n).astype(np.float32))
    spec = NetSpec("hh", [pop], rp.astype(np.uint32), post.astype(np.uint32), w,
                   np.ones(ns, dtype=np.uint8), dt, steps)
    if noise_sd > 0:
        spec.noise.append((0, np.full(n, noise_sd, np.float32), seed * 104729 + 3))
    if receptors:
        spec.w = np.abs(w).astype(np.float32)
        spec.receptors = dict(types=np.where(exc, 0, 2).astype(np.uint8), scale=float(g_unit))
    if stdp:
        spec.stdp = dict(a_plus=0.01, a_minus=0.012, tau_plus=20.0, tau_minus=20.0, w_min=0.0, w_max=1.0,
                         plastic=exc.astype(np.uint8))
    return spec


def build_device(spec: NetSpec, graph_steps: int = 8) -> CompiledNetwork:
    """Instantiate a spec on the device."""
    net = CompiledNetwork(spec.pops, spec.row_ptr, spec.col, spec.w, spec.delay, spec.dt)
    for s in spec.sources:
        net.add_poisson_source(s.rate_hz, s.dt_ms, s.seed, s.posts, s.weights, s.delays)
    for pop, sigma, seed in spec.noise:
        net.set_noise(pop, sigma, seed)
    if spec.stdp:
        st = dict(spec.stdp)
        plastic = st.pop("plastic", None)
        net.set_plasticity(plastic=plastic, **st)
    if spec.stp:
        net.set_short_term_plasticity(spec.stp["profiles"], spec.stp["dt"], spec.stp["syn_types"],
                                      spec.stp.get("src_types"))
    if spec.receptors:
        net.set_receptors(spec.receptors["types"], spec.receptors.get("scale", 1.0), spec.receptors.get("ext"))
    if spec.coupling:
        cp = spec.coupling
        if cp.get("gap") is not None:
            net.add_gap_junctions(*cp["gap"])
        if cp.get("graded") is not None:
            net.add_graded_synapses(*cp["graded"])
        for prof, conns in cp.get("sigmoid", []):
            pid = net.register_sigmoid_graded(*prof)
            net.add_sigmoid_graded(pid, *conns)
    net.set_graph_steps(graph_steps)
    return net


def gen_noise_seed(seed: int, fs: int) -> int:
    """Noise seed the device generator gives population `fs` (0 or 1)."""
    return (seed * 0x9E3779B97F4A7C15 + 17 + fs) & ((1 << 64) - 1)


def spec_from_generated(b, nid: int, g, steps: int) -> NetSpec:
    """A device-generated network of batch `b` as a NetSpec: CSR and
    populations downloaded, the per-neuron Poisson drive written as one
    explicit source per neuron (bit-identical, simuse_net_set_poisson_drive),
    the noise with the generator's seeds.  For checkers that rebuild the
    network on the host."""
    from . import _lib
    rp, col, w, d = b.download_csr(nid)
    pops = b.pops(nid)
    spec = NetSpec("generated", pops, rp, col, w, d, g.dt, steps)
    off = 0
    for p, pop in enumerate(pops):
        fs = 1 if (g.model == _lib.MODEL_IZHIKEVICH and p == 1) else 0
        if g.noise_sigma > 0:
            spec.noise.append((p, np.full(pop.n, g.noise_sigma, np.float32), gen_noise_seed(g.seed, fs)))
        if g.drive_rate_hz > 0:
            seed0 = (g.seed * 2654435761 + off) & 0xFFFFFFFF
            post = np.array([0], np.uint32)
            for i in range(pop.n):
                spec.sources.append(Source(g.drive_rate_hz, g.dt, (seed0 + i * 2654435761) & 0xFFFFFFFF,
                                           post + (off + i), np.array([g.drive_weight], np.float32),
                                           np.array([0], np.uint8)))
        off += pop.n
    return spec


def modular_small_world(n: int, modules: int = 64, k: int = 20, beta: float = 0.1, inter: float = 0.05,
                        coupling: float = 2.0, seed: int = 1):
    """Modular Watts–Strogatz graph for Kuramoto oscillators.

    n nodes in `modules` contiguous modules; inside a module each node links
    to its k/2 right neighbours on the module ring; each such edge is rewired
    with probability `inter` to a uniform node of another module, otherwise
    with probability `beta` to a uniform node of its own module.  Edges are
    made symmetric; W_ij = coupling / k_i (k_i = degree of i).  Returns
    (row_ptr, col, w, omega ~ N(0,1), theta0 ~ U(-pi, pi)).
    """
    rng = np.random.default_rng(seed)
    m = n // modules
    if m < k + 2:
        raise ValueError("modules too small for the lattice degree")
    sizes = np.full(modules, m, np.int64)
    sizes[: n - m * modules] += 1
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    node = np.arange(n, dtype=np.int64)
    mod = np.repeat(np.arange(modules), sizes)
    base = starts[mod]
    size = sizes[mod]
    local = node - base
    half = k // 2
    src = np.repeat(node, half)
    off = np.tile(np.arange(1, half + 1), n)
    dst = base.repeat(half) + (local.repeat(half) + off) % size.repeat(half)
    u = rng.random(src.size)
    inter_m = u < inter
    intra_m = (~inter_m) & (u < inter + beta)
    # rewire inside the module
    ni = int(intra_m.sum())
    s_mod = mod[src[intra_m]]
    r = rng.integers(0, sizes[s_mod] - 1)
    loc = local[src[intra_m]]
    r = r + (r >= loc)
    dst[intra_m] = starts[s_mod] + r
    # rewire to another module
    nx = int(inter_m.sum())
    om = (mod[src[inter_m]] + rng.integers(1, modules, size=nx)) % modules
    dst[inter_m] = starts[om] + rng.integers(0, sizes[om])
    a = np.concatenate([src, dst])
    b = np.concatenate([dst, src])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    deg = np.bincount(a, minlength=n)
    rp = np.zeros(n + 1, np.uint64)
    np.cumsum(deg, out=rp[1:])
    w = (coupling / deg[a]).astype(np.float32)
    omega = rng.normal(size=n).astype(np.float32)
    theta0 = rng.uniform(-np.pi, np.pi, n).astype(np.float32)
    return rp.astype(np.uint32), b.astype(np.uint32), w, omega, theta0


def _log_uniform(rng: np.random.Generator, lo: float, hi: float) -> int:
    return int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))


def _net_seed(master: int, k: int) -> int:
    z = (master * 0x9E3779B97F4A7C15 + (k + 1) * 0xD1B54A32D192ED03) & ((1 << 64) - 1)
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
    return z ^ (z >> 31)


def _gen_network(model: int, n: int, seed: int, rate_hz: float, k: int = 100):
    """One generated network of the batching and serving mixes: LIF (dt 0.1 ms,
    Brunel-style currents J*tau/dt: +20 / -100, delays 1..15), Izhikevich (dt 1 ms,
    exc U(0,0.5) / inh U(-1,0), delay 0, thalamic noise sd 5) or HH (dt 0.01 ms,
    delays 1..4, noise sd 30); every neuron has its own Poisson drive at rate_hz."""
    from . import _lib
    from .batch import gen_desc
    if model == _lib.MODEL_LIF:
        return gen_desc(model, n, k, seed, 0.1, w_exc=20.0, w_inh=100.0, d_min=1, d_max=15,
                        drive_rate_hz=rate_hz, drive_weight=20.0)
    if model == _lib.MODEL_IZHIKEVICH:
        return gen_desc(model, n, k, seed, 1.0, w_exc=0.5, w_inh=1.0, d_min=0, d_max=0,
                        drive_rate_hz=rate_hz, drive_weight=2.0, noise_sigma=5.0)
    return gen_desc(model, n, k, seed, 0.01, w_exc=2.0, w_inh=6.0, d_min=1, d_max=4,
                    drive_rate_hz=rate_hz, drive_weight=40.0, noise_sigma=30.0)


def batching_mix(count: int = 512, seed: int = 1, lo: int = 1_000, hi: int = 200_000, k: int = 100):
    """The batching configuration: `count` independent networks with sizes
    log-uniform in [lo, hi], model 50/50 LIF/Izhikevich, k synapses per neuron,
    per-network seed hash(seed, index) and Poisson drive rate U(1, 10) kHz."""
    from . import _lib
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = _log_uniform(rng, lo, hi)
        model = _lib.MODEL_LIF if rng.random() < 0.5 else _lib.MODEL_IZHIKEVICH
        out.append(_gen_network(model, n, _net_seed(seed, i), float(rng.uniform(1000.0, 10000.0)), k))
    return out


def serving_mix(count: int = 64, rate_hz: float = 200.0, seed: int = 1):
    """The serving configuration: Poisson arrivals at rate_hz; each request is a
    LIF, Izhikevich or HH network (5k-500k neurons, log-uniform) or a Kuramoto
    graph (10k-1M oscillators, 64 modules, Watts-Strogatz k=20, beta 0.1, 5%
    inter-module edges, K=2, RK4 at dt 1 ms), horizon U{500..20000} steps.
    Returns [(arrival time s, Request)]."""
    from . import _lib
    from .server import kuramoto_request, network_request
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate_hz))
        kind = int(rng.integers(0, 4))
        horizon = int(rng.integers(500, 20001))
        s = _net_seed(seed + 7, i)
        if kind == 3:
            n = _log_uniform(rng, 1e4, 1e6)
            out.append((t, kuramoto_request(n, modules=64, k=20, beta=0.1, inter=0.05, coupling=2.0, dt=0.001,
                                            seed=s, horizon=horizon)))
        else:
            model = (_lib.MODEL_LIF, _lib.MODEL_IZHIKEVICH, _lib.MODEL_HH)[kind]
            n = _log_uniform(rng, 5e3, 5e5)
            out.append((t, network_request(_gen_network(model, n, s, float(rng.uniform(1000.0, 10000.0))),
                                           horizon)))
    return out


# ---------------------------------------------------------------------------
# Creature pool (crates/dioma-engine/src/brain/dynamics/gpu_pool.rs)
# ---------------------------------------------------------------------------
@dataclass
class Creature:
    """A creature brain as ConnectomeData holds it
    (crates/SiMuse/crates/pillars-connectome/src/data.rs:80-103): CSR with local
    i32 indices, weights, and a transmitter sign per presynaptic neuron."""
    row_ptr: np.ndarray
    col: np.ndarray
    values: np.ndarray
    nt: np.ndarray

    @property
    def n(self) -> int:
        return int(self.row_ptr.size - 1)

    @property
    def n_syn(self) -> int:
        return int(self.col.size)


def creature(n: int, p: float = 0.08, seed: int = 1, w_lo: float = 10.0, w_hi: float = 60.0,
             inhib: float = 0.2) -> Creature:
    """Random recurrent creature: each ordered pair (i != j) connected with
    probability p, weights uniform in [w_lo, w_hi), a fraction ``inhib`` of the
    neurons inhibitory (nt_types -1)."""
    rng = np.random.default_rng(seed)
    adj = rng.random((n, n)) < p
    np.fill_diagonal(adj, False)
    pre, post = np.nonzero(adj)
    deg = np.bincount(pre, minlength=n)
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(deg, out=rp[1:])
    w = rng.uniform(w_lo, w_hi, size=post.size).astype(np.float32)
    nt = np.where(rng.random(n) < inhib, -1.0, 1.0).astype(np.float32)
    return Creature(rp, post.astype(np.int32), w, nt)


def pool_net_spec(creatures: List[Creature], delay: int, delays: Optional[List[np.ndarray]] = None,
                  config: Optional[dict] = None, steps: int = 0) -> NetSpec:
    """The pool's block-diagonal network in the engine's (push) form: one
    alpha-LIF population, weights signed by the presynaptic transmitter
    (fma(w, sign, acc) = acc + w*sign exactly), and delay 1 + max(d, 1) for a
    pool delay d: the pull at step s reads step s-1's spikes into slot
    s + max(d, 1), which the engine reads at step (s - 1) + 1 + max(d, 1)."""
    cfg = dict(dt_ms=0.5, tau_ref=2.2, tau_m=20.0, v_rest=-52.0, v_thresh=-45.0, v_reset=-52.0, tau_syn=3.0)
    if config:
        cfg.update({k: v for k, v in config.items() if k != "t_delay"})
    n = sum(c.n for c in creatures)
    pre_l, post_l, w_l, d_l = [], [], [], []
    off = 0
    for k, c in enumerate(creatures):
        deg = np.diff(c.row_ptr)
        pre = np.repeat(np.arange(c.n), deg) + off
        pre_l.append(pre)
        post_l.append(c.col.astype(np.int64) + off)
        sign = np.where(c.nt < 0, np.float32(-1.0), np.float32(1.0))
        w_l.append(c.values * sign[pre - off])
        dd = np.full(c.n_syn, delay, np.int64) if delays is None or delays[k] is None else delays[k].astype(np.int64)
        d_l.append(1 + np.maximum(dd, 1))
        off += c.n
    pre = np.concatenate(pre_l) if pre_l else np.zeros(0, np.int64)
    d = np.concatenate(d_l) if d_l else np.zeros(0, np.int64)
    if d.size and d.max() > 31:
        raise ValueError("pool_net_spec: a pool delay above 30 has no engine counterpart")
    rp = np.zeros(n + 1, np.uint32)
    np.cumsum(np.bincount(pre, minlength=n), out=rp[1:])
    col = np.concatenate(post_l).astype(np.uint32) if post_l else np.zeros(0, np.uint32)
    w = np.concatenate(w_l).astype(np.float32) if w_l else np.zeros(0, np.float32)
    pop = alpha_lif_pop(n, dt=cfg["dt_ms"], tau_ref=cfg["tau_ref"], tau_m=cfg["tau_m"], v_rest=cfg["v_rest"],
                        v_thresh=cfg["v_thresh"], v_reset=cfg["v_reset"], tau_syn=cfg["tau_syn"])
    return NetSpec("pool", [pop], rp, col, w, d.astype(np.uint8), float(cfg["dt_ms"]), steps)
