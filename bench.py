#!/usr/bin/env python3
"""Benchmark of the SiMuse network step on one MI355X.

Headline workload: the Izhikevich network of 1,000,000 neurons (800k regular
spiking, 200k fast spiking) with 100 random outgoing synapses each (100M
synapses, CSR, int32 indices), Gaussian thalamic input (sd 5 RS / 2 FS),
dt = 1 ms as two 0.5 ms half steps.  A "step" is one full network step:
synaptic input, neuron update, spike compaction, CSR push of every spike's
events.  Inputs are resident in HBM before the timed region, and the step
graph is captured and instantiated before it too.

Every delay is 0, so the network steps with one kernel per step,
megastep_kernel<Izhikevich> (the engine reports its step form): each block
gathers the previous step's events for its 2048 neurons, advances them and
writes its own spikes' events.  The roofline prices one launch at its
algorithmic bytes: the neuron-step bytes (v, u, ring-slot equivalent,
varying parameter columns) plus, per spiking neuron, its row bounds (8 B)
and, per synaptic event, target, delay, weight and the ring cell's read and
write (17 B) -- what the sequential CPU push moves.

Further legs: the Brunel LIF configuration on the GPU and on the CPU, the
batching configuration (512 networks, batched vs plain), the serving
configuration (Poisson arrivals, p50/p99 latency) and the Hodgkin-Huxley
precision configuration.  Each leg reports the FP32 result's error against
the oracle's FP64 twin (the checker) beside its throughput.

Prints one JSON line.  `value` is neuron-steps/s of the whole job;
synaptic events/s is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "dioma-core_amd"), os.path.join(REPO, "oracle")]

import numpy as np  # noqa: E402

METRIC = "synaptic events/s + neuron-steps/s (1 GPU); batched vs plain; p99 req latency"
HBM_PEAK_GBS = 8000.0
MEGA_THREADS, MEGA_NEURONS = 1024, 2048  # megastep block shape: threads, neurons (one bin)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads() -> int:
    """Host threads of this process's CPU share: the affinity mask, capped by
    OMP_NUM_THREADS when the environment sets it."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def izh_algorithmic_bytes_per_neuron(spec) -> float:
    """Bytes one Izhikevich neuron-step must move in this layout: v and u read
    and written (16 B), the ring slot read and zeroed (8 B), and 4 B for every
    parameter column that varies per neuron (uniform columns are scalars);
    the refractory timer is elided when tau_ref == 0."""
    total = 0.0
    for p in spec.pops:
        varying = 0
        for k in (1, 2, 3, 4, 6):  # a, b, c, d, v_thresh
            col = p.params[k, : p.n]
            if not np.all(col == col[0]):
                varying += 1
        refr = 0 if np.all(p.params[5, : p.n] <= 0) else 8
        total += p.n * (16 + 8 + 4 * varying + refr)
    return total / sum(p.n for p in spec.pops)


def hh_algorithmic_bytes_per_neuron(spec) -> float:
    """Bytes one Hodgkin-Huxley neuron-step must move: v, m, h, n and the
    refractory timer read and written (40 B; the CaT gate too, 8 B, where
    g_CaT != 0), the ring slot read and zeroed (8 B), each live receptor's
    conductance read and written (8 B each), both STDP traces read and
    written (16 B) when plasticity is on, and 4 B per varying parameter."""
    live = 0
    if spec.receptors is not None:
        types = np.asarray(spec.receptors["types"])
        live = len(set(int(t) for t in np.unique(types) if t < 7))
    total = 0.0
    for p in spec.pops:
        varying = sum(1 for k in range(1, 13) if not np.all(p.params[k, : p.n] == p.params[k, 0]))
        cat = 8 if np.any(p.params[10, : p.n] != 0) else 0
        total += p.n * (40 + cat + 8 + 8 * live + (16 if spec.stdp else 0) + 4 * varying)
    return total / sum(p.n for p in spec.pops)


def graph_steps_for(steps: int, most: int) -> int:
    """Steps per captured graph: the largest divisor of the timed step count
    up to `most` (at least 8 when one exists, else `most`), so that every timed
    step is a graph replay rather than a stand-alone launch."""
    best = max((d for d in range(1, min(64, max(most, 1)) + 1) if steps % d == 0), default=1)
    return best if best >= 8 else most


def form_name(form: int) -> str:
    import simuse
    return {simuse.FORM_ONE_KERNEL: "one_kernel", simuse.FORM_SPLIT: "split_two_kernel",
            simuse.FORM_BINNED: "binned_three_kernel", simuse.FORM_CHAINED: "chained"}[form]


# -- FP64 checks (the oracle's FP64 twin is the reference of every error) --------------
def fp64_window(net, spec, steps, threads) -> dict:
    """Start the oracle's FP64 twin from the device's own snapshot (every FP32
    value is exact in FP64), advance both `steps` steps and compare membrane
    potentials and spike counts: the error of that many FP32 steps from a
    common start."""
    import pyoracle as po
    snap = os.path.join(tempfile.mkdtemp(), "fp64_window.taas")
    net.save_state(snap)
    o64 = po.OracleNetwork(spec, 64, threads=threads)
    o64.set_state_from_snapshot(snap)
    os.remove(snap)
    s_dev0, s_ora0 = net.counters()[0], o64.counters()[0]
    out = {"steps": steps, "reference": "oracle FP64 twin started from the device snapshot"}
    done = 0
    for mark in sorted({min(10, steps), steps}):
        net.run(mark - done)
        o64.step(mark - done)
        done = mark
        dv = np.concatenate([np.abs(net.download_states(p)[0, : pp.n].astype(np.float64) - o64.state(p, 0))
                             for p, pp in enumerate(spec.pops)])
        out[f"v_mean_abs_err_mV_{mark}"] = float(dv.mean())
        out[f"v_max_abs_err_mV_{mark}"] = float(dv.max())
        out[f"frac_over_1mV_{mark}"] = float(np.mean(dv > 1.0))
    s_dev, s_ora = net.counters()[0] - s_dev0, o64.counters()[0] - s_ora0
    out["spikes_fp32"], out["spikes_fp64"] = int(s_dev), int(s_ora)
    out["spike_count_rel_err"] = abs(s_dev - s_ora) / max(s_ora, 1)
    o64.close()
    return out


def cpu_baseline(spec, steps: int, threads: int, warmup: int = 10):
    """The oracle's FP32 path (the reference's Network::step restated: threaded
    population loop and push_spikes_par) on the host; returns neuron-steps/s,
    synaptic events/s and the wall time of the timed steps."""
    import pyoracle as po
    o = po.OracleNetwork(spec, 32, threads=threads)
    o.step(warmup)
    e0 = o.counters()[1]
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    ev = o.counters()[1] - e0
    o.close()
    return spec.n_total * steps / dt, ev / dt, dt


# -- Izhikevich 1M / 100M: the headline ------------------------------------------------
def bench_izhikevich(args) -> dict:
    import simuse
    from simuse.workloads import build_device, izhikevich_network

    log(f"building Izhikevich network N={args.neurons} k={args.synapses_per_neuron}")
    t0 = time.perf_counter()
    spec = izhikevich_network(n=args.neurons, k=args.synapses_per_neuron, seed=1,
                              steps=args.steps + args.warmup)
    log(f"  generated in {time.perf_counter() - t0:.1f}s")
    gsteps = graph_steps_for(args.steps, args.graph_steps)
    net = build_device(spec, graph_steps=gsteps)
    form = net.step_form()
    log(f"  device bytes {net.device_bytes() / 2**30:.2f} GiB, step form {form_name(form)}")
    tuned = None
    if form != simuse.FORM_ONE_KERNEL:
        # tune the two-kernel forms' schedule on the steady state, not on the transient
        net.run(max(1, args.warmup // 2))
        tuned = {"scatter": "tile" if net.autotune(32) == simuse.SCATTER_TILE else "wave",
                 "block_tiles": net.block_tiles()}
    net.run(args.warmup)
    # one untimed replay of the step graph, whatever --warmup is, through the
    # same timing call as the region: the graph is captured, instantiated and
    # launched, and the timing path's own first-call costs are paid, before
    # the timed region
    net.time_steps(gsteps)
    s0, e0 = net.counters()
    net.sync()
    # The timed region: the K steps between two HIP events on the network's
    # stream, bracketed by synchronisations.
    t0 = time.perf_counter()
    region_ms = net.time_steps(args.steps)
    net.sync()
    wall = time.perf_counter() - t0
    s1, e1 = net.counters()
    spikes, events = s1 - s0, e1 - e0
    ms_per_step = wall * 1e3 / args.steps
    log(f"  {ms_per_step * 1e3:.2f} us/step (events {region_ms * 1e3 / args.steps:.2f} us), "
        f"{spikes / args.steps:.0f} spikes/step, {events / args.steps:.0f} events/step")

    # The step kernels' device time: one more replay of the step graph queued
    # behind a device-side spacer, between two HIP events on the network's
    # stream (one kernel per step in the one-kernel form)
    qs0, qe0 = net.counters()
    q_ms = net.time_steps_queued(gsteps)
    qs1, qe1 = net.counters()

    # per-phase HIP events around each kernel (no graph) over profile_steps more
    net.set_profiling(True)
    ps0, pe0 = net.counters()
    net.run(args.profile_steps)
    k_ms = net.kernel_times()
    ps1, pe1 = net.counters()
    net.set_profiling(False)
    step_ms, stdp_ms, insert_ms, fixup_ms = k_ms
    bpn = izh_algorithmic_bytes_per_neuron(spec)
    neuron_bytes = bpn * spec.n_total
    # The step kernel's average launch duration: the queued replay's event time
    # per step in the one-kernel form (one launch per step), HIP events around
    # each step kernel otherwise; algorithmic bytes of those same steps.
    one_kernel = form == simuse.FORM_ONE_KERNEL
    if one_kernel:
        kernel_ms = q_ms / gsteps
        pev, psp = (qe1 - qe0) / gsteps, (qs1 - qs0) / gsteps
    else:
        kernel_ms = step_ms
        pev, psp = (pe1 - pe0) / args.profile_steps, (ps1 - ps0) / args.profile_steps
    kernel_bytes = neuron_bytes + pev * 17 + psp * 8  # the step kernel also pushes its spikes' events
    achieved = kernel_bytes / (kernel_ms * 1e-3) / 1e9
    step_bytes = neuron_bytes + (events / args.steps) * 17 + (spikes / args.steps) * 8
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_izhikevich_1m.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("kernel_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    fp64 = None
    if not args.no_fp64:
        log(f"  FP64 window: {args.fp64_steps} steps from the device state")
        fp64 = fp64_window(net, spec, args.fp64_steps, host_threads())
        log(f"    |dv| after {args.fp64_steps} steps: mean {fp64[f'v_mean_abs_err_mV_{args.fp64_steps}']:.2e} mV, "
            f">1 mV {fp64[f'frac_over_1mV_{args.fp64_steps}']:.2%}, spikes {fp64['spikes_fp32']} vs "
            f"{fp64['spikes_fp64']}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        log(f"cpu baseline: oracle FP32 path, {args.cpu_steps} steps on {threads} threads, "
            f"{max(4, args.cpu_steps // 8)} on 1")
        v_all, ev_all, t_all = cpu_baseline(spec, args.cpu_steps, threads)
        v_one, ev_one, t_one = cpu_baseline(spec, max(4, args.cpu_steps // 8), 1)
        cpu = {"value": v_all, "unit": "neuron-steps/s", "cores": threads, "kind": "port",
               "sample": f"same 1M/100M network from its initial state, {args.cpu_steps} steps after 10 warm-up "
                         f"steps ({t_all:.1f}s) on {threads} threads (host CPU share; nproc "
                         f"{os.cpu_count()}): threaded population loop, threaded noise and push_spikes_par; "
                         f"1 thread: {v_one:.3e} neuron-steps/s over {max(4, args.cpu_steps // 8)} steps "
                         f"({t_one:.1f}s)",
               "synaptic_events_per_s": ev_all, "one_core_value": v_one, "one_core_synaptic_events_per_s": ev_one}
        log(f"  {v_all:.3e} neuron-steps/s on {threads} threads, {v_one:.3e} on 1")
    shape = ({"kernel": "megastep_kernel<Izhikevich>", "threads_per_block": MEGA_THREADS,
              "neurons_per_block": MEGA_NEURONS, "blocks": -(-spec.n_total // MEGA_NEURONS)} if one_kernel
             else tuned)
    out = {
        "metric": METRIC,
        "value": spec.n_total * args.steps / wall,
        "unit": "neuron-steps/s",
        "synaptic_events_per_s": events / wall,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Izhikevich 2003 network, random-init parameters and weights)",
        "config": {"workload": "izhikevich_1M_100M", "neurons": spec.n_total,
                   "synapses": spec.n_syn, "dt_ms": 1.0, "graph_steps": gsteps,
                   "step_form": form_name(form), "shape": shape,
                   "untimed_steps_before_region": args.warmup + gsteps,
                   "spikes_per_step": spikes / args.steps, "events_per_step": events / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "megastep_kernel<Izhikevich> (input gather, update, compaction, push)",
                     "algorithmic_bytes": kernel_bytes,
                     "algorithmic_bytes_per_neuron": bpn, "algorithmic_bytes_per_event": 17,
                     "algorithmic_bytes_per_spike": 8,
                     "kernel_ms": kernel_ms,
                     "kernel_ms_source": (f"HIP events on the network's stream around one replay of the "
                                          f"{gsteps}-step graph right after the timed region, queued behind a "
                                          f"device-side spacer (no host submission gap), per step = per launch; "
                                          f"algorithmic bytes from those steps' events" if one_kernel else
                                          f"HIP events around each step kernel over {args.profile_steps} "
                                          f"profiled steps"),
                     "region_ms_per_step": region_ms / args.steps},
        # HIP-event phases without graphs: step kernels (megastep, or neuron
        # update with its fused push), STDP, stand-alone push writers, apply /
        # fix-up; an empty phase reads the event gap
        "kernels_ms_per_step": {"step_kernel": step_ms, "stdp": stdp_ms, "push_writers": insert_ms,
                                "apply_fixup": fixup_ms, "steps": args.profile_steps,
                                "events_per_step": (pe1 - pe0) / args.profile_steps,
                                "spikes_per_step": (ps1 - ps0) / args.profile_steps},
        "step_frac_of_peak": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "fp64_error": fp64,
        "cpu_baseline": cpu,
    }
    net.close()
    return out


# -- Brunel LIF 10k on the GPU and on the CPU ------------------------------------------
def bench_brunel(args) -> dict:
    """The Brunel LIF configuration (10k neurons, p = 0.01, delays 1..15,
    2 kHz Poisson per neuron, dt 0.1 ms) over its full 10,000 steps on the
    device and on the host (1 thread and the CPU share), with the FP64 twin's
    firing statistics beside the device's."""
    import pyoracle as po
    from simuse import stats
    from simuse.workloads import brunel_lif_network, build_device
    T = args.c1_steps
    spec = brunel_lif_network(n=10_000, p=0.01, seed=1, steps=T, rate_hz=2000.0)
    log(f"brunel: LIF {spec.n_total} neurons, {spec.n_syn} synapses, {len(spec.sources)} Poisson sources, {T} steps")
    gsteps = graph_steps_for(T, args.graph_steps)
    net = build_device(spec, graph_steps=gsteps)
    form = net.step_form()
    net.run(gsteps)  # graph built and replayed once; counters below are differences
    s0, e0 = net.counters()
    net.sync()
    t0 = time.perf_counter()
    region_ms = net.time_steps(T)
    net.sync()
    wall = time.perf_counter() - t0
    s1, e1 = net.counters()
    net.close()
    gpu = {"steps": T, "wall_s": wall, "us_per_step": wall * 1e6 / T, "region_us_per_step": region_ms * 1e3 / T,
           "neuron_steps_per_s": spec.n_total * T / wall, "synaptic_events_per_s": (e1 - e0) / wall,
           "spikes_per_step": (s1 - s0) / T, "events_per_step": (e1 - e0) / T, "step_form": form_name(form)}
    log(f"  GPU {gpu['us_per_step']:.2f} us/step, {gpu['spikes_per_step']:.2f} spikes/step, "
        f"{gpu['events_per_step']:.0f} events/step")
    threads = host_threads()
    v1, ev1, t1 = cpu_baseline(spec, T, 1)
    va, eva, ta = cpu_baseline(spec, T, threads)
    cpu = {"one_core": {"neuron_steps_per_s": v1, "synaptic_events_per_s": ev1, "wall_s": t1},
           "all_cores": {"threads": threads, "neuron_steps_per_s": va, "synaptic_events_per_s": eva, "wall_s": ta},
           "kind": "port", "nproc": os.cpu_count()}
    log(f"  CPU {v1:.3e} neuron-steps/s on 1 thread, {va:.3e} on {threads}")
    fp64 = None
    if not args.no_fp64:
        # statistics of the device run vs the FP64 twin over the whole horizon
        dev = build_device(spec, graph_steps=gsteps)
        dev.record(T)
        dev.run(T)
        rd = dev.download_raster()
        dev.close()
        o64 = po.OracleNetwork(spec, 64, threads=threads)
        o64.record(True)
        o64.step(T)
        ro = o64.raster().reshape(T, -1)
        o64.close()
        sd, so = stats.summary(rd, spec.dt * 1e-3), stats.summary(ro, spec.dt * 1e-3)
        fp64 = {"reference": "oracle FP64 twin, same network and inputs, whole horizon", "fp32": sd, "fp64": so,
                **stats.rel_errors(sd, so)}
        log(f"  rate {sd['rate_hz']:.4f} vs FP64 {so['rate_hz']:.4f} Hz")
    return {"workload": "Brunel LIF 10k neurons (80/20), p=0.01, J_E +0.1 mV / J_I -0.5 mV, delays 1-15 steps, "
                        f"2 kHz Poisson per neuron, dt 0.1 ms, {T} steps",
            "synapses": spec.n_syn, "gpu": gpu, "cpu": cpu, "fp64_error": fp64,
            "note": "under the stated 2 kHz drive (about 4 mV mean depolarisation against a 15 mV gap to "
                    "threshold) the network is nearly silent: the step is launch- and latency-bound"}


# -- 512 uneven networks batched -------------------------------------------------------
def bench_batching(args) -> dict:
    """Batching configuration: 512 generated networks (log-uniform 1k-200k
    neurons, LIF/Izhikevich, 100 synapses each, Poisson drive U(1,10) kHz)
    stepped together for `batch_steps` steps in 16-step graphs, against plain
    per-network runs (one network per engine, same graphs) on a random sample.
    networks/s = networks simulated for the full horizon per second of stepping
    (network construction excluded on both sides)."""
    import pyoracle as po
    import simuse
    from simuse import stats
    from simuse.batch import Batch
    from simuse.workloads import batching_mix, spec_from_generated
    descs = batching_mix(args.batch_networks, seed=1)
    T = args.batch_steps
    log(f"batching: {len(descs)} networks, sum N {sum(d.n for d in descs)}, {T} steps")
    t0 = time.perf_counter()
    b = Batch()
    ids = [b.add_generated(d) for d in descs]
    b.set_graph_steps(16)
    b.run(1)  # finalize every network
    built = time.perf_counter() - t0
    mode = b.autotune(16)  # schedule, block size and step form, timed on the live state
    one_kernel = b.one_kernel()
    b.run(16)
    c0 = [b.counters(i) for i in ids]
    b.sync()
    t0 = time.perf_counter()
    b.step_async(T)
    b.sync()
    wall = time.perf_counter() - t0
    c1 = [b.counters(i) for i in ids]
    ev = sum(y[1] - x[1] for x, y in zip(c0, c1))
    nsteps = sum(d.n for d in descs) * T
    b.close()
    batched = {"networks": len(descs), "steps": T, "wall_s": wall, "networks_per_s": len(descs) / wall,
               "neuron_steps_per_s": nsteps / wall, "synaptic_events_per_s": ev / wall,
               "scatter": "tile" if mode == simuse.SCATTER_TILE else "wave",
               "step_form": "one_kernel" if one_kernel else "binned", "build_s": built}
    log(f"  batched: {wall:.3f} s for {T} steps, {len(descs) / wall:.1f} networks/s")
    rng = np.random.default_rng(2)
    sample = rng.choice(len(descs), size=min(args.plain_sample, len(descs)), replace=False)
    plain_wall = 0.0
    plain_nsteps = 0
    for i in sample:
        d = descs[int(i)]
        pb = Batch()
        pb.add_generated(d)
        pb.set_graph_steps(16)
        pb.run(17)
        t0 = time.perf_counter()
        pb.step_async(T)
        pb.sync()
        plain_wall += time.perf_counter() - t0
        plain_nsteps += d.n * T
        pb.close()
    plain = {"sample_networks": int(len(sample)), "wall_s": plain_wall,
             "networks_per_s": len(sample) / plain_wall, "neuron_steps_per_s": plain_nsteps / plain_wall}
    log(f"  plain: {len(sample)} networks one at a time in {plain_wall:.3f} s, "
        f"{len(sample) / plain_wall:.1f} networks/s")

    fp64 = None
    if not args.no_fp64:
        # a sample of the mix (two LIF, two Izhikevich among the networks of at
        # most 20k neurons) rerun alone with recording: spike counts equal to
        # the batched run's, statistics against the FP64 twin
        small = [i for i, d in enumerate(descs) if d.n <= 20_000]
        pick = [i for i in small if descs[i].model == simuse.MODEL_LIF][:2] + \
               [i for i in small if descs[i].model == simuse.MODEL_IZHIKEVICH][:2]
        rows = []
        for i in pick:
            d = descs[i]
            ab = Batch()
            nid = ab.add_generated(d)
            spec = spec_from_generated(ab, nid, d, T)  # the generator's initial state
            ab.set_graph_steps(16)
            ab.record(nid, T + 17)
            ab.run(T + 17)
            rd = ab.download_raster(nid)
            ab.close()
            # the batched run timed steps [17, 17 + T) of every network
            alone = int(rd[17: 17 + T].sum())
            o64 = po.OracleNetwork(spec, 64, threads=host_threads())
            o64.record(True)
            o64.step(T)
            ro = o64.raster().reshape(T, -1)
            o64.close()
            dt_s = d.dt * 1e-3
            sd, so = stats.summary(rd[:T], dt_s), stats.summary(ro, dt_s)
            rows.append({"network": i, "model": "lif" if d.model == simuse.MODEL_LIF else "izhikevich", "n": d.n,
                         "spikes_alone_equal_batched": alone == c1[i][0] - c0[i][0],
                         "fp32": sd, "fp64": so, **stats.rel_errors(sd, so)})
            log(f"  fp64 net {i} ({rows[-1]['model']}, n={d.n}): rate {sd['rate_hz']:.3f}/{so['rate_hz']:.3f} "
                f"cv {sd['isi_cv']:.3f}/{so['isi_cv']:.3f} chi {sd['chi']:.4f}/{so['chi']:.4f}, alone == batched "
                f"{rows[-1]['spikes_alone_equal_batched']}")
        fp64 = {"reference": f"oracle FP64 twin, {T} steps from the generated state", "sample": rows}
    return {"workload": f"{len(descs)} generated networks, 1k-200k neurons log-uniform, LIF/Izhikevich 50/50, "
                        f"100 syn/neuron, Poisson drive U(1,10) kHz, {T} steps",
            "batched": batched, "plain": plain,
            "batched_vs_plain": batched["networks_per_s"] / plain["networks_per_s"], "fp64_error": fp64}


# -- serving a Poisson request stream --------------------------------------------------
def bench_serving(args, budget_gb: float, fp64: bool = True) -> dict:
    """Serving configuration: Poisson arrivals (200/s) of LIF/Izhikevich/HH
    networks (5k-500k neurons) and Kuramoto graphs (10k-1M oscillators, RK4),
    horizons 500-20,000 steps, admitted by an HBM budget and stepped together
    (continuous batching).  Latency = completion - submission."""
    from simuse.server import DONE, Server
    from simuse.workloads import serving_mix
    reqs = serving_mix(args.serve_requests, rate_hz=args.serve_rate, seed=3)
    log(f"serving: {len(reqs)} requests at {args.serve_rate}/s, budget {budget_gb} GB")
    srv = Server(int(budget_gb * (1 << 30)), chunk_steps=16)
    ids = []
    t0 = time.perf_counter()
    for t, r in reqs:
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        ids.append(srv.submit(r))
    last = time.perf_counter()
    while True:
        st = srv.stats()
        if st.completed >= st.submitted or time.perf_counter() - t0 > args.serve_timeout:
            break
        time.sleep(0.05)
        if time.perf_counter() - last > 5.0:
            log(f"  t={time.perf_counter() - t0:.1f}s completed {st.completed}/{st.submitted}, live {st.live}, "
                f"queued {st.queued}")
            last = time.perf_counter()
    srv.drain(5.0)
    wall = time.perf_counter() - t0
    res = [srv.result(i) for i in ids]
    st = srv.stats()
    srv.close()
    done = [x for x in res if x.status == DONE]
    lat = np.array([x.t_done - x.t_submit for x in done])
    wait = np.array([x.t_admit - x.t_submit for x in done])
    span = max(x.t_done for x in done) - min(x.t_submit for x in res) if done else float("nan")
    out = {"workload": f"{len(reqs)} requests, Poisson arrivals {args.serve_rate}/s, LIF/Izh/HH 5k-500k neurons "
                       "or Kuramoto 10k-1M (RK4), horizons 500-20000 steps",
           "completed": len(done), "rejected_or_failed": len(res) - len(done),
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "mean_admission_wait_s": float(wait.mean()), "max_admission_wait_s": float(wait.max()),
           "networks_per_s": len(done) / span,
           "wall_s": wall, "peak_live": int(st.peak_live), "peak_bytes": int(st.peak_bytes),
           "budget_bytes": int(budget_gb * (1 << 30)),
           "admitted_after_a_completion": int(sum(1 for x in done if x.t_admit > min(y.t_done for y in done)))}
    log(f"  {len(done)} done, p50 {out['p50_latency_s']:.3f} s, p99 {out['p99_latency_s']:.3f} s, "
        f"{out['networks_per_s']:.1f} networks/s, peak live {st.peak_live}, admission wait max "
        f"{out['max_admission_wait_s']:.3f} s")
    if fp64 and not args.no_fp64:
        out["fp64_error"] = serving_fp64(reqs, res, args)
    return out


def serving_fp64(reqs, res, args) -> dict:
    """FP32 vs FP64 for a sample of served requests: the two network requests
    with the least work (neurons x horizon) get their spike count over the
    horizon compared with the FP64 twin's; the Kuramoto request with the least
    work gets its order parameter R compared with the FP64 twin's at
    min(horizon, kura_fp64_steps) steps, on the graph rebuilt alone."""
    import pyoracle as po
    import simuse
    from simuse.batch import Batch
    from simuse.kuramoto import KuramotoGraph
    from simuse.server import DONE, REQ_KURAMOTO
    from simuse.workloads import spec_from_generated
    nets = [(r.net.n * r.horizon, k) for k, (_, r) in enumerate(reqs) if r.kind != REQ_KURAMOTO]
    kuras = [(r.kura.n * r.horizon, k) for k, (_, r) in enumerate(reqs) if r.kind == REQ_KURAMOTO]
    rows = []
    for _, k in sorted(nets)[:2]:
        r = reqs[k][1]
        if res[k].status != DONE:
            continue
        b = Batch()
        nid = b.add_generated(r.net)
        spec = spec_from_generated(b, nid, r.net, int(r.horizon))
        b.close()
        o64 = po.OracleNetwork(spec, 64, threads=host_threads())
        o64.step(int(r.horizon))
        s64 = o64.counters()[0]
        o64.close()
        s32 = int(res[k].spikes)
        model = {simuse.MODEL_LIF: "lif", simuse.MODEL_IZHIKEVICH: "izhikevich", simuse.MODEL_HH: "hh"}[r.net.model]
        rows.append({"request": k, "model": model, "n": int(r.net.n), "horizon": int(r.horizon),
                     "spikes_fp32": s32, "spikes_fp64": int(s64),
                     "rate_rel_err": abs(s32 - s64) / max(s64, 1)})
        log(f"  fp64 request {k} ({model}, n={r.net.n}, {r.horizon} steps): spikes {s32} vs {s64}")
    if kuras:
        _, k = sorted(kuras)[0]
        r = reqs[k][1]
        T = int(min(r.horizon, args.kura_fp64_steps))
        g = KuramotoGraph.generated(r.kura, "rk4")
        rp, col, w, om = g.download_graph()
        th0 = g.phases()
        g.run(T)
        r32 = g.order_parameter()
        g.close()
        ok = po.OracleKuramoto(rp, col, w, om, th0.astype(np.float32), r.kura.dt)
        ok.rk4(T)
        r64 = ok.order()
        rows.append({"request": k, "model": "kuramoto", "n": int(r.kura.n), "steps": T,
                     "served_horizon": int(r.horizon), "R_fp32": r32, "R_fp64": r64, "R_abs_err": abs(r32 - r64)})
        log(f"  fp64 request {k} (Kuramoto n={r.kura.n}, {T} steps): R {r32:.6f} vs {r64:.6f}")
    return {"reference": "oracle FP64 twin, same generated networks and inputs", "sample": rows}


# -- HH 200k precision network ---------------------------------------------------------
def bench_precision(args) -> dict:
    """Precision configuration: the Hodgkin-Huxley network with heavy-tailed
    out-degree, lognormal conductance synapses (AMPA from excitatory, GABA_A
    from inhibitory neurons) and pair STDP on its excitatory synapses
    (dt = 0.01 ms).  Every window of 1000 steps starts the oracle's FP64
    twin from the device's own state (its `.taas` snapshot: every FP32 value
    is exact in FP64), so each window's error is that of 1000 FP32 steps
    from a common start, on a reduced network (the oracle runs on the host
    cores); the device's step time is taken at the full 200k-neuron size."""
    import pyoracle as po
    from simuse.workloads import build_device, hh_network
    n, windows = args.prec_neurons, args.prec_windows
    spec = hh_network(n=n, seed=3, steps=windows * 1000, noise_sd=30.0, receptors=True)
    log(f"precision: HH {n} neurons, {spec.n_syn} conductance synapses, pair STDP, {windows} windows of 1000 steps "
        "vs FP64")
    dev = build_device(spec, graph_steps=16)
    o64 = po.OracleNetwork(spec, 64, threads=host_threads())
    plastic = spec.stdp["plastic"].astype(bool)
    rows = []
    snap = os.path.join(tempfile.mkdtemp(), "window.taas")
    for k in range(windows):
        if k:  # the FP64 twin restarts from the device's state
            dev.save_state(snap)
            o64.set_state_from_snapshot(snap)
        s_dev0, s_ora0 = dev.counters()[0], o64.counters()[0]
        dev.run(100)  # 1 ms in: the FP32 rounding error before spike-time flips spread
        o64.step(100)
        dv100 = np.abs(dev.download_states(0)[0, :n].astype(np.float64) - o64.state(0, 0))
        dev.run(900)
        o64.step(900)
        v_dev = dev.download_states(0)[0, :n].astype(np.float64)
        v64 = o64.state(0, 0)
        dv = np.abs(v_dev - v64)
        g_dev = dev.download_conductances().astype(np.float64)
        g64 = o64.conductances()
        dg = np.abs(g_dev - g64)
        w_dev = dev.download_weights().astype(np.float64)[plastic]
        w64 = o64.weights()[plastic]
        dw = np.abs(w_dev - w64)
        s_dev, s_ora = dev.counters()[0], o64.counters()[0]
        rate_dev = (s_dev - s_dev0) / n / (1000 * spec.dt * 1e-3)
        rate_ora = (s_ora - s_ora0) / n / (1000 * spec.dt * 1e-3)
        rows.append({"window": k, "v_mean_abs_err_mV_100": float(dv100.mean()),
                     "v_max_abs_err_mV_100": float(dv100.max()),
                     "v_mean_abs_err_mV": float(dv.mean()), "v_max_abs_err_mV": float(dv.max()),
                     "frac_over_1mV": float(np.mean(dv > 1.0)),
                     "g_mean_abs_err": float(dg.mean()), "g_max_abs_err": float(dg.max()),
                     "w_rms_err": float(np.sqrt(np.mean(dw ** 2))), "w_max_abs_err": float(dw.max()),
                     "rate_hz_fp32": float(rate_dev), "rate_hz_fp64": float(rate_ora),
                     "rate_rel_err": float(abs(rate_dev - rate_ora) / max(rate_ora, 1e-12))})
        log(f"  window {k}: 100 steps |dv| mean {dv100.mean():.2e} max {dv100.max():.2e} mV; 1000 steps |dv| mean "
            f"{dv.mean():.2e} max {dv.max():.2e} mV, >1 mV {np.mean(dv > 1.0):.2%}, |dg| max {dg.max():.2e}, "
            f"w rms {rows[-1]['w_rms_err']:.2e}, rate {rate_dev:.2f} vs {rate_ora:.2f} Hz")
    dev.close()
    big = hh_network(n=args.prec_full_neurons, seed=3, steps=max(args.prec_horizon, args.prec_full_steps + 208),
                     noise_sd=30.0, receptors=True)
    net = build_device(big, graph_steps=16)
    form = net.step_form()
    net.run(208)
    net.sync()
    s0, e0 = net.counters()
    t0 = time.perf_counter()
    ms = net.time_steps(args.prec_full_steps)
    net.sync()
    us = (time.perf_counter() - t0) * 1e6 / args.prec_full_steps
    s1, e1 = net.counters()
    # the rest of the stated 50,000-step horizon on the device (its FP64
    # comparison over the whole horizon runs on a 10k-neuron network of the
    # same construction in the configuration tests)
    horizon = {}
    done = 208 + args.prec_full_steps
    if args.prec_horizon > done:
        plastic = big.stdp["plastic"].astype(bool)
        w0 = float(big.w.astype(np.float64)[plastic].mean())
        t1 = time.perf_counter()
        net.run(args.prec_horizon - done)
        net.sync()
        secs = time.perf_counter() - t1
        s2, e2 = net.counters()
        w = net.download_weights().astype(np.float64)[plastic]
        horizon = {"steps": args.prec_horizon, "simulated_ms": args.prec_horizon * big.dt,
                   "device_s_for_remaining_steps": secs, "remaining_steps": args.prec_horizon - done,
                   "us_per_step": secs * 1e6 / (args.prec_horizon - done),
                   "spikes": s2, "events": e2,
                   "rate_hz": s2 / big.n_total / (args.prec_horizon * big.dt * 1e-3),
                   "mean_plastic_weight": float(w.mean()), "mean_plastic_weight_at_start": w0,
                   "plastic_weights_at_bounds": float(np.mean((w <= 0.0) | (w >= 1.0)))}
        log(f"  full horizon {args.prec_horizon} steps: {secs:.2f} s for the last {args.prec_horizon - done}, "
            f"rate {horizon['rate_hz']:.2f} Hz, mean plastic weight {w0:.5f} -> {w.mean():.5f}")
    net.close()
    q_ms = ms / args.prec_full_steps  # device time per step: the region's events over many graph replays
    sp, ev = (s1 - s0) / args.prec_full_steps, (e1 - e0) / args.prec_full_steps
    bpn = hh_algorithmic_bytes_per_neuron(big)
    step_bytes = bpn * big.n_total + ev * 17 + sp * 8
    achieved = step_bytes / (q_ms * 1e-3) / 1e9
    log(f"  full size {args.prec_full_neurons} neurons / {big.n_syn} synapses: {us:.2f} us/step wall, "
        f"{q_ms * 1e3:.2f} us/step device, {sp:.0f} spikes and {ev:.0f} events per step, "
        f"{achieved:.0f} GB/s algorithmic")
    return {"workload": f"HH {n} neurons ({spec.n_syn} synapses), Pareto out-degree, lognormal conductances "
                        "(AMPA excitatory, GABA_A inhibitory, 0.003 mS/cm2 units), pair STDP (A+ 0.01, A- 0.012, "
                        "tau 20 ms, [0,1]) on excitatory synapses, Gaussian input sd 30, dt 0.01 ms",
            "fp64_error": {"reference": "oracle FP64 twin, restarted from the device state at every window",
                           "windows": rows},
            "full_size": {"neurons": args.prec_full_neurons, "synapses": big.n_syn, "step_form": form_name(form),
                          "steps": args.prec_full_steps, "us_per_step": us,
                          "region_us_per_step": ms * 1e3 / args.prec_full_steps,
                          "spikes_per_step": sp, "events_per_step": ev, "full_horizon": horizon,
                          "roofline": {"bound": "hbm", "kernel": step_kernels(form),
                                       "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                                       "algorithmic_bytes_per_step": step_bytes,
                                       "algorithmic_bytes_per_neuron": bpn, "algorithmic_bytes_per_event": 17,
                                       "algorithmic_bytes_per_spike": 8,
                                       "note": "STDP weight traffic (FP64 masters of the spiking rows' edges) "
                                               "left out: the bytes are a lower bound",
                                       "step_ms": q_ms,
                                       "step_ms_source": "HIP events around the timed region (graph replays) "
                                                         "/ steps"}}}


def step_kernels(form: int) -> str:
    """The launches of one step of a plastic receptor network in `form`."""
    import simuse
    if form == simuse.FORM_SPLIT:
        return ("one step: front_kernel<HH, receptors> (the bin's conductance jumps and currents in event order, "
                "update, receptor currents, STDP traces, compaction), bin_kernel (STDP, event records, in-edge LTP)")
    return ("one step: neuron_kernel<HH> (update, receptor currents, STDP traces, compaction), bin_kernel (STDP, "
            "event records), bin_apply_kernel<receptors>")


# -- Kuramoto: the phase-coupling kernel at 1M oscillators / 20M edges -----------------
def kura_algorithmic_bytes(n: int, n_edges: int, row_weights: bool = False) -> dict:
    """Bytes an RK4 step must move at least once per stage: per edge its
    column (4 B), weight (4 B) and the neighbour's (sin, cos) (8 B); per node
    the row pointer (4), own (sin, cos) (8), omega (4), FP64 phase (8), the
    FP64 RK4 accumulator (written in stage 1, read and written in 2 and 3,
    read in 4), the new (sin, cos) (8) and the phase written back in stage 4.
    A graph whose rows each carry one weight and share one degree (the
    modular small world) moves 12 B per edge: the row's weight (4 B per node)
    takes the row pointer's place."""
    per_edge = 12 if row_weights else 16
    node = {1: 4 + 8 + 4 + 8 + 8 + 8, 2: 4 + 8 + 4 + 8 + 16 + 8, 3: 4 + 8 + 4 + 8 + 16 + 8,
            4: 4 + 8 + 4 + 8 + 8 + 8 + 8}
    return {"per_edge_per_stage": per_edge, "per_node_per_step": sum(node.values()),
            "per_step": 4 * per_edge * n_edges + sum(node.values()) * n}


def bench_kuramoto(args) -> dict:
    """RK4 Kuramoto on a modular small-world graph built on the device:
    1M oscillators in 64 modules, 20 in-edges each (20M edges), W = 2/20,
    dt 0.01.  Timed with HIP events per step (four stage kernels) after a
    warm-up; R is reported with the FP64-phase order parameter."""
    from simuse.kuramoto import KuramotoGraph
    from simuse.server import KuraGenDesc
    desc = KuraGenDesc(args.kura_nodes, 64, args.kura_k, 0.1, 0.05, 2.0, 0.01, 7)
    g = KuramotoGraph.generated(desc, "rk4")
    log(f"kuramoto: {g.n} oscillators, {g.n_edges} edges, RK4")
    g.run(20)
    g.set_profiling(True)
    g.run(args.kura_steps)
    step_ms = g.step_ms()
    g.set_profiling(False)
    g.sync()
    t0 = time.perf_counter()
    g.run(args.kura_steps)
    wall_us = (time.perf_counter() - t0) * 1e6 / args.kura_steps
    r = g.order_parameter()
    g.close()
    ab = kura_algorithmic_bytes(g.n, g.n_edges, row_weights=True)  # K / k on every edge of a row, k per row
    achieved = ab["per_step"] / (step_ms * 1e-3) / 1e9
    # L2-miss bytes of the same workload's stage launches (FETCH_SIZE and
    # WRITE_SIZE passes, copy-calibrated), per RK4 step
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_kuramoto_1m_20m.json")
    if g.n == 1_000_000 and args.kura_k == 20 and os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_step")
        except (OSError, ValueError):
            traffic = None
    log(f"  {step_ms * 1e3:.1f} us/step (HIP events, 4 stages), {wall_us:.1f} us/step wall, "
        f"{achieved:.0f} GB/s algorithmic, R {r:.4f}")
    return {"workload": f"Kuramoto RK4, {g.n} oscillators, {g.n_edges} edges (64 modules, k {args.kura_k}, "
                        "beta 0.1, inter 0.05), W = 2/k, dt 0.01",
            "steps": args.kura_steps, "us_per_step": wall_us, "order_parameter": r,
            "oscillator_steps_per_s": g.n * 1e6 / wall_us,
            "roofline": {"bound": "hbm", "kernel": "kura_stage_kernel (4 launches per RK4 step)",
                         "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_per": "RK4 step (4 launches)",
                         "algorithmic_bytes_per_step": ab["per_step"],
                         "algorithmic_bytes_per_edge_per_stage": ab["per_edge_per_stage"],
                         "algorithmic_bytes_per_node_per_step": ab["per_node_per_step"],
                         "kernel_ms_per_step": step_ms,
                         "kernel_ms_source": f"HIP events around the 4 stage kernels of each of {args.kura_steps} "
                                             "steps (no graph)"}}


# -- creature pool: many small alpha-LIF brains, per-post pull scatter -----------------
# -- every neuron model as a 32M-neuron population (the single-population path) ------
# State columns each model's update reads and writes, and parameter columns it
# reads, in the single-population layout (every parameter per neuron, dt from
# the call; HH's CaT gate only where g_CaT != 0, here 0)
MODEL_COLUMNS = {"lif": (2, 5), "izhikevich": (3, 6), "hh": (5, 12), "gpea": (9, 13), "alpha_lif": (3, 6)}


def bench_models(args) -> dict:
    """Each neuron model on its own through the single-population API (the
    CompiledModel path): --model-neurons neurons with a resident random input
    current, one launch of the per-neuron update per step.  Device time
    (HIP events behind a spacer) against the bytes the update must move:
    state columns read and written, parameter columns read, the input read
    and the spike flag written; the oracle's FP32 path on the host cores
    beside it."""
    import pyoracle as po
    from simuse.model import CompiledModel
    from simuse.models import GPEA_PULSE, alpha_lif_pop, gpea_pop, hh_pop, izh_pop, lif_pop
    from simuse.workloads import NetSpec
    n, steps = args.model_neurons, args.model_steps
    rng = np.random.default_rng(5)
    cases = [
        ("lif", lif_pop(n, 0.1, 0.05, -65.0, -50.0, -65.0, 2.0, rng.uniform(-65, -50, n).astype(np.float32)),
         0.1, (0.0, 30.0)),
        ("izhikevich", izh_pop(n, 1.0, 0.02, 0.2, -65.0, 8.0, 0.0, 30.0, -65.0, -13.0), 1.0, (0.0, 15.0)),
        ("hh", hh_pop(n, 0.01, v0=rng.uniform(-70, -60, n).astype(np.float32)), 0.01, (0.0, 20.0)),
        ("gpea", gpea_pop(n), 0.025, (0.0, 3.0 * GPEA_PULSE)),
        ("alpha_lif", alpha_lif_pop(n), 0.5, (0.0, 200.0)),
    ]
    threads = host_threads()
    out = {"neurons": n, "steps": steps, "path": "simuse_pop_* (pop_step_kernel, one thread per neuron)",
           "note": "at the default 32M neurons the working sets (1.4-4.2 GB) are far above the MI355X's 256 MB "
                   "Infinity Cache; at 8M it still serves part of the light models' bytes (LIF 0.92 of the HBM "
                   "peak), at 1M nearly all (LIF 1.005)",
           "legs": {}}
    for name, pop, dt, (lo, hi) in cases:
        cur = rng.uniform(lo, hi, n).astype(np.float32)
        m = CompiledModel(pop.model, n, pop.stride)
        m.upload(pop.states, pop.params, cur)
        m.step_graph(dt, 20)
        m.sync()
        ms = m.time_steps(dt, steps)
        spikes = int(m.download_spikes().sum())
        m.close()
        ns_rw, np_r = MODEL_COLUMNS[name]
        bpn = 8 * ns_rw + 4 * np_r + 4 + 4
        step_ms = ms / steps
        achieved = bpn * n / (step_ms * 1e-3) / 1e9
        # the oracle: the same population as a network without synapses with
        # the current injected every step, about a second of host time
        rp = np.zeros(n + 1, np.uint32)
        spec = NetSpec(f"pop_{name}", [pop], rp, np.zeros(0, np.uint32), np.zeros(0, np.float32),
                       np.zeros(0, np.uint8), dt, 0)
        o = po.OracleNetwork(spec, 32, threads=threads)
        idx = np.arange(n, dtype=np.uint32)
        o.inject_current(0, idx, cur)
        t1 = time.perf_counter()
        o.step()
        per = max(time.perf_counter() - t1, 1e-6)
        cs = int(min(100, max(3, 1.0 / per)))
        t1 = time.perf_counter()
        for _ in range(cs):
            o.inject_current(0, idx, cur)
            o.step()
        cpu = n * cs / (time.perf_counter() - t1)
        o.close()
        out["legs"][name] = {
            "us_per_step": step_ms * 1e3, "neuron_steps_per_s": n / (step_ms * 1e-3), "spikes_last_step": spikes,
            "roofline": {"bound": "hbm", "kernel": f"pop_step_kernel<{name}>", "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "algorithmic_bytes_per_neuron": bpn, "kernel_ms": step_ms,
                         "kernel_ms_source": "HIP events around the timed launches, queued behind a spacer / steps"},
            "cpu_baseline": {"value": cpu, "unit": "neuron-steps/s", "cores": threads, "kind": "port",
                             "sample": f"{cs} steps of the same population (current injected each step) on the "
                                       "oracle's FP32 path"}}
        log(f"  model {name}: {step_ms * 1e3:.2f} us/step on the device, {bpn} B/neuron -> {achieved:.0f} GB/s "
            f"({achieved / HBM_PEAK_GBS:.3f}); CPU {cpu:.3e} neuron-steps/s on {threads} threads")
    return out


def bench_pool(args) -> dict:
    """The game engine's creature pool (BrainPool): `pool_creatures` random
    brains of `pool_neurons` alpha-LIF neurons (connection probability
    0.02, 20% inhibitory), a tenth of each brain's neurons receiving a voltage
    input each tick, 100 substeps per tick (two submits, 64 + 36).  Timed:
    whole ticks with the spike-history readback the reference does (wall),
    and the substeps' device time without readback (HIP events).  The first
    tick's history is checked against the oracle pool; the CPU baseline is
    the oracle pool (single thread) on one tick."""
    import pyoracle as po
    from simuse.pool import BrainPool
    from simuse.workloads import creature
    nc, nn, spt = args.pool_creatures, args.pool_neurons, 100
    rng = np.random.default_rng(5)
    cs = [creature(nn, p=0.02, seed=1000 + k) for k in range(nc)]
    pool = BrainPool()
    for c in cs:
        pool.register(c.row_ptr, c.col, c.values, c.nt)
    n_total, n_syn, _, delay = pool.info()
    log(f"pool: {nc} creatures x {nn} alpha-LIF neurons = {n_total}, {n_syn} synapses, {spt} substeps per tick")
    exts_all = []
    for t in range(args.pool_ticks + 1):
        ex = []
        for c in cs:
            e = np.zeros(c.n, np.float32)
            m = rng.random(c.n) < 0.1
            e[m] = rng.uniform(0.0, 12.0, int(m.sum())).astype(np.float32)
            ex.append(e)
        exts_all.append(ex)
    # parity of the first tick against the oracle pool (also the CPU baseline)
    ora = po.OraclePool()
    for c in cs:
        ora.register(c.row_ptr, c.col, c.values, c.nt)
    flat0, _ = pool.step(exts_all[0], spt)
    t0 = time.perf_counter()
    ref0 = ora.step(exts_all[0], spt)
    cpu_s = time.perf_counter() - t0
    ora.close()
    same = bool(np.array_equal(flat0, ref0))
    outdeg = np.concatenate([np.diff(c.row_ptr) for c in cs]).astype(np.int64)
    pool.sync_weights([(c.values, True) for c in cs])  # the reference's per-tick weight upload
    # whole ticks with readback
    spikes = act_edges = 0
    t0 = time.perf_counter()
    for t in range(1, args.pool_ticks + 1):
        flat, _ = pool.step(exts_all[t], spt)
        f = flat.reshape(spt, n_total)
        spikes += int(f.sum())
        act_edges += int((f.astype(np.int64) @ outdeg).sum())
    wall = time.perf_counter() - t0
    # device time of the substeps alone (same input, no readback)
    dev_ms = pool.time_substeps(args.pool_ticks * spt)
    pool.close()
    subs = args.pool_ticks * spt
    us_sub = dev_ms * 1e3 / subs
    per_sub = n_total * 40 + n_syn * 4 + (act_edges / subs) * 4
    achieved = per_sub / (us_sub * 1e-6) / 1e9
    log(f"  {wall * 1e3 / args.pool_ticks:.2f} ms per tick with readback, {us_sub:.2f} us per substep on the device, "
        f"{spikes / subs:.0f} spikes per substep, {achieved:.0f} GB/s algorithmic; first tick == oracle: {same}; "
        f"oracle tick {cpu_s:.2f} s")
    return {"workload": f"creature pool: {nc} alpha-LIF brains x {nn} neurons ({n_total}), {n_syn} synapses "
                        f"(p 0.02, 20% inhibitory), delay {delay}, 100 substeps per tick, voltage input on a tenth",
            "ticks": args.pool_ticks, "ms_per_tick_with_readback": wall * 1e3 / args.pool_ticks,
            "ticks_per_s": args.pool_ticks / wall, "readback_bytes_per_tick": spt * n_total * 4,
            "device_us_per_substep": us_sub, "neuron_substeps_per_s_device": n_total / (us_sub * 1e-6),
            "spikes_per_substep": spikes / subs, "first_tick_equals_oracle": same,
            "roofline": {"bound": "hbm", "kernel": "pool_tick_kernel (one launch per submit chunk of up to 64 "
                                                   "substeps, one workgroup per creature: update + pull scatter "
                                                   "per substep, substeps separated by block barriers)",
                         "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "algorithmic_bytes_per_substep": per_sub,
                         "algorithmic_bytes": "40 B per neuron (v, g, refr_cnt read and written, slot read and "
                                              "zeroed, history, row pointer), 4 B per in-edge (source and bucket), "
                                              "4 B per in-edge whose source spiked (signed weight); the ring adds "
                                              "left out (lower bound). The one-launch chunk keeps state, ring and "
                                              "in-edges on chip between substeps, so these are the bytes a substep's "
                                              "work reads and writes, not HBM traffic",
                         "kernel_ms_per_substep": dev_ms / subs},
            "cpu_baseline": {"value": n_total * spt / cpu_s, "unit": "neuron-substeps/s", "cores": 1,
                             "kind": "port", "sample": f"oracle pool (restated per-post pull), first tick, {spt} "
                                                       "substeps, single thread"}}


def copy_calibration(mib: int = 1024) -> int:
    """Device-to-device copy of a known size (the HIP runtime's copy kernel),
    used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device copies of 1 GiB (traffic counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--synapses-per-neuron", type=int, default=100)
    ap.add_argument("--graph-steps", type=int, default=64,
                    help="most steps per captured graph (the bench uses the largest divisor of --steps up to it)")
    ap.add_argument("--profile-steps", type=int, default=100,
                    help="steps timed per kernel with HIP events for the phase breakdown (at least 1)")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-twin error checks")
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--only-headline", action="store_true", help="run the Izhikevich leg only")
    ap.add_argument("--no-brunel", action="store_true", help="skip the Brunel LIF leg")
    ap.add_argument("--c1-steps", type=int, default=10_000)
    ap.add_argument("--no-batching", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batch-networks", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-gb", type=float, default=160.0)
    ap.add_argument("--serve-tight-gb", type=float, default=6.0,
                    help="budget of the second serving run, small enough that requests wait for memory")
    ap.add_argument("--serve-timeout", type=float, default=300.0)
    ap.add_argument("--kura-fp64-steps", type=int, default=2000)
    ap.add_argument("--no-kuramoto", action="store_true", help="skip the Kuramoto leg")
    ap.add_argument("--only-kuramoto", action="store_true", help="run the Kuramoto leg only (profiling)")
    ap.add_argument("--kura-nodes", type=int, default=1_000_000)
    ap.add_argument("--kura-k", type=int, default=20)
    ap.add_argument("--kura-steps", type=int, default=100)
    ap.add_argument("--no-pool", action="store_true", help="skip the creature-pool leg")
    ap.add_argument("--no-models", action="store_true", help="skip the per-model population leg")
    ap.add_argument("--only-models", action="store_true", help="run the per-model population leg only")
    ap.add_argument("--model-neurons", type=int, default=32_000_000,
                    help="population size of the per-model leg (32M: 1.4-4.2 GB working sets, far above the "
                         "256 MB Infinity Cache, so the bytes come from HBM)")
    ap.add_argument("--model-steps", type=int, default=50)
    ap.add_argument("--only-pool", action="store_true", help="run the creature-pool leg only (profiling)")
    ap.add_argument("--pool-creatures", type=int, default=64)
    ap.add_argument("--pool-neurons", type=int, default=1000)
    ap.add_argument("--pool-ticks", type=int, default=20)
    ap.add_argument("--no-precision", action="store_true", help="skip the HH precision leg")
    ap.add_argument("--only-precision", action="store_true", help="run the HH precision leg only (profiling)")
    ap.add_argument("--prec-neurons", type=int, default=10_000)
    ap.add_argument("--prec-windows", type=int, default=5)
    ap.add_argument("--prec-full-neurons", type=int, default=200_000)
    ap.add_argument("--prec-full-steps", type=int, default=2000)
    ap.add_argument("--prec-horizon", type=int, default=50_000,
                    help="steps of the full-size precision network on the device (its stated horizon)")
    args = ap.parse_args()
    if args.profile_steps < 1:
        ap.error("--profile-steps must be at least 1")
    if args.copy_calibration:
        return copy_calibration()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(a single network fits one MI355X; independent networks shard by network)", file=sys.stderr)
        return 2

    import simuse
    if simuse.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1
    if args.only_kuramoto:
        print(json.dumps({"kuramoto": bench_kuramoto(args)}), flush=True)
        return 0
    if args.only_precision:
        print(json.dumps({"precision": bench_precision(args)}), flush=True)
        return 0
    if args.only_models:
        print(json.dumps({"models": bench_models(args)}), flush=True)
        return 0
    if args.only_pool:
        print(json.dumps({"pool": bench_pool(args)}), flush=True)
        return 0
    out = bench_izhikevich(args)
    if not args.only_headline:
        if not args.no_brunel:
            out["brunel"] = bench_brunel(args)
        if not args.no_batching:
            out["batching"] = bench_batching(args)
        if not args.no_serving:
            out["serving"] = bench_serving(args, args.serve_budget_gb)
            out["serving_tight_budget"] = bench_serving(args, args.serve_tight_gb, fp64=False)
        if not args.no_kuramoto:
            out["kuramoto"] = bench_kuramoto(args)
        if not args.no_pool:
            out["pool"] = bench_pool(args)
        if not args.no_models:
            out["models"] = bench_models(args)
        if not args.no_precision:
            out["precision"] = bench_precision(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
