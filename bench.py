#!/usr/bin/env python3
"""Benchmark of the SiMuse network step on one MI355X.

Workload: the Izhikevich network of 1,000,000 neurons (800k regular spiking,
200k fast spiking) with 100 random outgoing synapses each (100M synapses,
CSR, int32 indices), Gaussian thalamic input (sd 5 RS / 2 FS), dt = 1 ms as
two 0.5 ms half steps.  A "step" is one full network step: synaptic input,
neuron update, spike compaction, CSR push of every spike's events.  Inputs
are resident in HBM before the timed region.

The whole step is one kernel, megastep_kernel<Izhikevich> (every delay is
0, so the network runs the binned scatter in its one-kernel form): each
block gathers the previous step's events for its 2048 neurons, advances
them and writes its own spikes' events.  Its algorithmic bytes per launch
are the neuron-step bytes (v, u, ring slot, varying parameter columns) plus,
per spiking neuron, its row bounds (8 B) and, per synaptic event, target,
delay, weight and the ring cell's read and write (17 B) -- what the
sequential CPU push moves.

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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def one_kernel_step(spec) -> bool:
    """The engine steps a network with one kernel launch per step when every
    delay is at most one step, all populations share one model, there is no
    plasticity or receptor synapse, and its bins (2048 neurons each) number at
    most 2048."""
    import numpy as np
    return (int(np.max(spec.delay, initial=0)) <= 1 and len({p.model for p in spec.pops}) == 1
            and not spec.stdp and not getattr(spec, "receptors", None) and spec.n_total <= 2048 * 2048)


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


def cpu_baseline(spec, steps: int, threads: int):
    import pyoracle as po
    o = po.OracleNetwork(spec, 32, threads=threads)
    o.step(3)  # warm-up
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return spec.n_total * steps / dt, dt


def bench_batching(args) -> dict:
    """Batching configuration: 512 generated networks (log-uniform 1k-200k
    neurons, LIF/Izhikevich, 100 synapses each, Poisson drive U(1,10) kHz)
    stepped together for `batch_steps` steps in 16-step graphs, against plain
    per-network runs (one network per engine, same graphs) on a random sample.
    networks/s = networks simulated for the full horizon per second of stepping
    (network construction excluded on both sides)."""
    import simuse
    from simuse.batch import Batch
    from simuse.workloads import batching_mix
    descs = batching_mix(args.batch_networks, seed=1)
    T = args.batch_steps
    log(f"batching: {len(descs)} networks, sum N {sum(d.n for d in descs)}, {T} steps")
    t0 = time.perf_counter()
    b = Batch()
    ids = [b.add_generated(d) for d in descs]
    b.set_graph_steps(16)
    b.run(1)  # finalize every network
    built = time.perf_counter() - t0
    mode = b.autotune(16)
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
               "scatter": "tile" if mode == simuse.SCATTER_TILE else "wave", "build_s": built}
    log(f"  batched: {wall:.3f} s for {T} steps, {len(descs) / wall:.1f} networks/s")
    rng = np.random.default_rng(2)
    sample = rng.choice(len(descs), size=min(args.plain_sample, len(descs)), replace=False)
    plain_wall = 0.0
    plain_nsteps = 0
    for i in sample:
        d = descs[int(i)]
        pb = Batch()
        nid = pb.add_generated(d)
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
    return {"workload": f"{len(descs)} generated networks, 1k-200k neurons log-uniform, LIF/Izhikevich 50/50, "
                        f"100 syn/neuron, Poisson drive U(1,10) kHz, {T} steps",
            "batched": batched, "plain": plain,
            "batched_vs_plain": batched["networks_per_s"] / plain["networks_per_s"]}


def bench_serving(args) -> dict:
    """Serving configuration: Poisson arrivals (200/s) of LIF/Izhikevich/HH
    networks (5k-500k neurons) and Kuramoto graphs (10k-1M oscillators, RK4),
    horizons 500-20,000 steps, admitted by an HBM budget and stepped together
    (continuous batching).  Latency = completion - submission."""
    from simuse.server import DONE, Server
    from simuse.workloads import serving_mix
    reqs = serving_mix(args.serve_requests, rate_hz=args.serve_rate, seed=3)
    log(f"serving: {len(reqs)} requests at {args.serve_rate}/s, budget {args.serve_budget_gb} GB")
    srv = Server(int(args.serve_budget_gb * (1 << 30)), chunk_steps=16)
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
           "mean_admission_wait_s": float(wait.mean()), "networks_per_s": len(done) / span,
           "wall_s": wall, "peak_live": int(st.peak_live), "peak_bytes": int(st.peak_bytes),
           "budget_bytes": int(args.serve_budget_gb * (1 << 30))}
    log(f"  {len(done)} done, p50 {out['p50_latency_s']:.3f} s, p99 {out['p99_latency_s']:.3f} s, "
        f"{out['networks_per_s']:.1f} networks/s, peak live {st.peak_live}")
    return out


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
    o64 = po.OracleNetwork(spec, 64, threads=min(16, os.cpu_count() or 1))
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
    big = hh_network(n=args.prec_full_neurons, seed=3, steps=args.prec_full_steps + 200, noise_sd=30.0,
                     receptors=True)
    net = build_device(big, graph_steps=16)
    net.run(200)
    net.sync()
    t0 = time.perf_counter()
    net.step_graph(args.prec_full_steps)
    net.sync()
    us = (time.perf_counter() - t0) * 1e6 / args.prec_full_steps
    spikes, events = net.counters()
    net.close()
    log(f"  full size {args.prec_full_neurons} neurons / {big.n_syn} synapses: {us:.2f} us/step")
    return {"workload": f"HH {n} neurons ({spec.n_syn} synapses), Pareto out-degree, lognormal conductances "
                        "(AMPA excitatory, GABA_A inhibitory, 0.003 mS/cm2 units), pair STDP (A+ 0.01, A- 0.012, "
                        "tau 20 ms, [0,1]) on excitatory synapses, Gaussian input sd 30, dt 0.01 ms",
            "reference": "oracle FP64 twin, restarted from the device state at every window",
            "windows": rows,
            "full_size": {"neurons": args.prec_full_neurons, "synapses": big.n_syn,
                          "steps": args.prec_full_steps, "us_per_step": us}}


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
    ap.add_argument("--graph-steps", type=int, default=16)
    ap.add_argument("--profile-steps", type=int, default=100,
                    help="steps timed per kernel with HIP events for the roofline (at least 1)")
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scatter", choices=["auto", "tile", "wave"], default="auto")
    ap.add_argument("--no-batching", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batch-networks", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-gb", type=float, default=160.0)
    ap.add_argument("--serve-timeout", type=float, default=300.0)
    ap.add_argument("--no-precision", action="store_true", help="skip the HH precision leg")
    ap.add_argument("--prec-neurons", type=int, default=10_000)
    ap.add_argument("--prec-windows", type=int, default=5)
    ap.add_argument("--prec-full-neurons", type=int, default=200_000)
    ap.add_argument("--prec-full-steps", type=int, default=2000)
    args = ap.parse_args()
    if args.profile_steps < 1:
        ap.error("--profile-steps must be at least 1 (the roofline needs the step kernel's time)")
    if args.copy_calibration:
        return copy_calibration()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(a single network fits one MI355X; independent networks shard by network)", file=sys.stderr)
        return 2

    import simuse
    from simuse.workloads import build_device, izhikevich_network

    if simuse.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    log(f"building Izhikevich network N={args.neurons} k={args.synapses_per_neuron}")
    t0 = time.perf_counter()
    spec = izhikevich_network(n=args.neurons, k=args.synapses_per_neuron, seed=1,
                              steps=args.steps + args.warmup)
    log(f"  generated in {time.perf_counter() - t0:.1f}s")
    net = build_device(spec, graph_steps=args.graph_steps)
    log(f"  device bytes {net.device_bytes() / 2**30:.2f} GiB")
    if args.scatter == "auto":
        # tune on the network's steady state, not on the initial transient
        net.run(max(1, args.warmup // 2))
        mode = net.autotune(32)
    else:
        net.set_scatter_mode(simuse.SCATTER_TILE if args.scatter == "tile" else simuse.SCATTER_WAVE)
        mode = net.scatter_mode()
    log(f"  scatter schedule {'tile' if mode == simuse.SCATTER_TILE else 'wave'}, "
        f"{net.block_tiles()} tiles per block")

    net.run(args.warmup)
    s0, e0 = net.counters()
    net.sync()
    # The timed region: the K steps between two HIP events on the network's
    # stream (one megastep launch per step), bracketed by synchronisations.
    t0 = time.perf_counter()
    region_ms = net.time_steps(args.steps)
    net.sync()
    wall = time.perf_counter() - t0
    s1, e1 = net.counters()
    spikes = s1 - s0
    events = e1 - e0
    ms_per_step = wall * 1e3 / args.steps
    neuron_steps_per_s = spec.n_total * args.steps / wall
    events_per_s = events / wall
    log(f"  {ms_per_step * 1e3:.2f} us/step, {spikes / args.steps:.0f} spikes/step, "
        f"{events / args.steps:.0f} events/step")

    # Live per-kernel durations: HIP events around each kernel on the
    # network's stream, over profile_steps further steps.
    net.set_profiling(True)
    ps0, pe0 = net.counters()
    net.run(args.profile_steps)
    k_ms = net.kernel_times()
    ps1, pe1 = net.counters()
    net.set_profiling(False)
    neuron_ms, stdp_ms, insert_ms, fixup_ms = k_ms
    bpn = izh_algorithmic_bytes_per_neuron(spec)
    neuron_bytes = bpn * spec.n_total
    ev_step = (pe1 - pe0) / max(1, args.profile_steps)
    act_step = (ps1 - ps0) / max(1, args.profile_steps)
    scatter_bytes = ev_step * 17 + act_step * 8
    kernel_bytes = neuron_bytes + scatter_bytes  # the step kernel also pushes its spikes' events
    # the step kernel's average launch duration over the timed region: the
    # network runs one kernel per step, so the region's event time per step
    # (launch gaps included) bounds it from above
    kernel_ms = region_ms / args.steps if one_kernel_step(spec) else neuron_ms
    achieved = kernel_bytes / (kernel_ms * 1e-3) / 1e9
    step_bytes = kernel_bytes
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_izhikevich_1m.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("kernel_bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle FP32 path, {args.cpu_steps} steps on {threads} threads")
        v_all, t_all = cpu_baseline(spec, args.cpu_steps, threads)
        v_one, t_one = cpu_baseline(spec, max(4, args.cpu_steps // 8), 1)
        cpu = {"value": v_all, "unit": "neuron-steps/s", "cores": threads, "kind": "port",
               "sample": f"same 1M/100M network, {args.cpu_steps} steps after 3 warm-up steps "
                         f"({t_all:.1f}s); 1 thread: {v_one:.3e} neuron-steps/s over "
                         f"{max(4, args.cpu_steps // 8)} steps ({t_one:.1f}s)",
               "one_core_value": v_one}

    out = {
        "metric": METRIC,
        "value": neuron_steps_per_s,
        "unit": "neuron-steps/s",
        "synaptic_events_per_s": events_per_s,
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
                   "synapses": spec.n_syn, "dt_ms": 1.0, "graph_steps": args.graph_steps,
                   "scatter": "tile" if mode == simuse.SCATTER_TILE else "wave",
                   "block_tiles": net.block_tiles(),
                   "spikes_per_step": spikes / args.steps, "events_per_step": events / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "megastep_kernel<Izhikevich> (input gather, update, compaction, push)",
                     "algorithmic_bytes": kernel_bytes,
                     "algorithmic_bytes_per_neuron": bpn, "algorithmic_bytes_per_event": 17,
                     "algorithmic_bytes_per_spike": 8,
                     "kernel_ms": kernel_ms,
                     "kernel_ms_source": ("HIP events over the timed region / steps (one kernel per step)"
                                          if one_kernel_step(spec) else "HIP events around the step kernel")},
        # HIP-event phases: step kernels (megastep, or neuron update with its
        # fused push), STDP, stand-alone push writers, apply / fix-up; an empty
        # phase reads the ~4.5 us event gap
        "kernels_ms_per_step": {"step_kernel": neuron_ms, "stdp": stdp_ms, "push_writers": insert_ms,
                                "apply_fixup": fixup_ms},
        "step_frac_of_peak": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "cpu_baseline": cpu,
    }
    net.close()
    if not args.no_batching:
        out["batching"] = bench_batching(args)
    if not args.no_serving:
        out["serving"] = bench_serving(args)
    if not args.no_precision:
        out["precision"] = bench_precision(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
