#!/usr/bin/env python3
"""Benchmark of the exchange-step hot path on one MI355X.

Workload (the plain drop-in): the reference-size cerebellar scaffold, 96,737
cells and ~3.2M synapses from the seeded scaffold generator (88,158 granule,
7,073 glomeruli fed by 2,400 mossy fibres, 69 Purkinje, 219 Golgi, 603 stellate,
603 basket, 12 DCN), FP32 state, FP32 synaptic multipliers, 4 Hz mossy background
plus a 100 Hz burst to a 50-fibre patch at 300-350 ms.  A step is one exchange
step (1 ms of simulated time = 10 Euler sub-steps of 0.1 ms).

Prints one JSON line: synaptic events/s (value) and real-time factor of the timed
steps, the roofline figure of the dominant kernel (the gather kernel) measured
with HIP events, the CPU baseline (the single-threaded C oracle on a bounded
sample of the same network), and the batched BF16 form (64 uneven instances
advanced together).
"""
import argparse
import json
import os
import re
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "synaptic events/s + real-time factor, 96k scaffold; batched BF16 vs plain FP32"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_env(pkg, preset, plasticities=0, **over):
    env = pkg.NetworkEnvironment(1, 7, 16, 2, plasticities)
    cfg = pkg.synth_preset(preset)
    for k, v in over.items():
        setattr(cfg, k, v)
    env.describe_synthetic(cfg)
    return env


def advance(env, start, n, chunk=25):
    for s in range(start, start + n, chunk):
        env.loop_n_steps(s, min(chunk, start + n - s), 0, want_output=False)


def outdegree(env):
    info = env.info()
    deg = np.zeros(info.total_neurons, np.int64)
    for i in range(info.n_projections):
        p = env.projection(i)
        if p.pre < 0:
            continue
        _, col, _ = env.projection_csr(i)
        np.add.at(deg, env.population(p.pre).base_id + col.astype(np.int64), 1)
    return deg


def cpu_baseline(env, steps, threads):
    """The oracle (plain C restatement of the reference step) on the first `steps`
    exchange steps of the same network: once on one thread (the reference runs
    its CPU path as one pthread per simulation, GPUScaffoldSimulation.cu:7-53),
    once with OpenMP over `threads` host cores (propagation and neuron update
    split by post neuron)."""
    import oracle
    deg = outdegree(env)

    def timed(nthreads):
        ref = oracle.OracleSim.from_env(env)
        ref.set_threads(nthreads)
        ref.initialize()
        spent = 0.0
        events = 0
        for s in range(steps):
            t0 = time.perf_counter()
            ref.step(s)
            spent += time.perf_counter() - t0
            events += int(deg[ref.spike_row(s).astype(bool)].sum())
        return events, spent

    ev1, t1 = timed(1)
    evn, tn = timed(threads)
    assert ev1 == evn
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"host_cpu": model, "host_cpus_visible": os.cpu_count(), "value": ev1 / t1, "unit": "synaptic events/s", "cores": 1, "kind": "port",
            "sample": f"first {steps} exchange steps ({steps} ms simulated) of the same 96.7k-cell network, "
                      f"{t1:.1f} s on one thread",
            "real_time_factor": steps * 1e-3 / t1, "ms_per_step": t1 * 1e3 / steps,
            "all_cores": {"value": evn / tn, "unit": "synaptic events/s", "cores": threads,
                          "real_time_factor": steps * 1e-3 / tn, "ms_per_step": tn * 1e3 / steps,
                          "sample": f"same {steps} steps, OpenMP over {threads} threads, {tn:.1f} s"}}


def batched(pkg, n_inst, steps, warmup, threads):
    """64 independent instances of uneven size (0.25x..4x of the 96.7k scaffold),
    each with its own seed and mossy rate U(1, 60) Hz, BF16 multipliers, advanced
    together by one step kernel per exchange step (hipGraph replays)."""
    scales = ([0.25] * 16 + [0.5] * 16 + [1.0] * 16 + [2.0] * 8 + [4.0] * 8)[:n_inst]
    rng = np.random.default_rng(1234)
    rates = rng.uniform(1.0, 60.0, len(scales))
    envs = [None] * len(scales)
    errs = []

    def work(idx):
        try:
            for i in idx:
                e = pkg.NetworkEnvironment(1, 7, 16, 2, 0)
                cfg = pkg.synth_preset(1)
                cfg.scale = scales[i]
                cfg.seed = 1000 + i
                cfg.stim_seed = 5000 + i
                cfg.mf_rate_hz = float(rates[i])
                cfg.burst_n_mf = max(1, int(round(50 * scales[i])))
                e.describe_synthetic(cfg)
                e.set_option("weight_precision", pkg.WEIGHTS_BF16)
                e.set_option("neuron_seed", 77 + i)
                for k, v in layout.items():
                    e.set_option(k, v)
                e.initialize(50, 25)
                envs[i] = e
        except Exception as ex:  # surfaced below
            errs.append(ex)

    # autotune the layout on one instance of the largest scale (4x; the 4x
    # instances hold 42 % of the batch's synapses, and at that size the step is
    # throughput-bound like the batch) and use it for all
    probe = pkg.NetworkEnvironment(1, 7, 16, 2, 0)
    cfg = pkg.synth_preset(1)
    cfg.scale = max(scales)
    cfg.burst_n_mf = max(1, int(round(50 * cfg.scale)))
    cfg.mf_rate_hz = float(np.median(rates))
    probe.describe_synthetic(cfg)
    probe.set_option("weight_precision", pkg.WEIGHTS_BF16)
    tuned = probe.autotune(100)
    probe.finalize()
    layout = {k: tuned[k] for k in ("lane_max_fanin", "lane_synapses", "graph_steps", "index_bits", "stage_spikes")}

    order = sorted(range(len(scales)), key=lambda i: -scales[i])
    parts = [order[k::threads] for k in range(threads)]
    ths = [threading.Thread(target=work, args=(p,)) for p in parts]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    if errs:
        raise errs[0]
    b = pkg.Batch(envs)  # replays graphs of the first instance's graph_steps
    b.run(warmup)
    before = sum(e.totals()[1] for e in envs)
    t0 = time.perf_counter()
    ms = b.time(steps)
    wall = time.perf_counter() - t0
    after = sum(e.totals()[1] for e in envs)
    syn = sum(e.info().total_synapses for e in envs)
    neurons = sum(e.info().total_neurons for e in envs)
    b.destroy()
    for e in envs:
        e.finalize()
    return {"value": (after - before) / wall, "unit": "synaptic events/s", "instances": len(scales),
            "synapses_scanned_per_s": syn * steps / wall,
            "neurons": int(neurons), "synapses": int(syn), "steps": steps, "ms_per_step": wall * 1e3 / steps,
            "gpu_ms_per_step": ms / steps, "dtype": "bf16 multipliers, f32 state",
            "aggregate_real_time_factor": len(scales) * steps * 1e-3 / wall, "autotuned": tuned}


def plain_variant(pkg, steps, warmup, plastic=False, **options):
    """The plain FP32 96.7k scaffold with other storage options (e.g. 16-bit
    column indices: same results bit for bit, other memory traffic), or with the
    reference's parallel-fibre-to-Purkinje Teacher rule (A1_J 0, A2_JI -0.01,
    window 50; SynapticPlasticity.cu:10-16), basket-to-Purkinje standing in for
    the climbing fibre as the teacher projection."""
    env = make_env(pkg, 1, plasticities=1 if plastic else 0)
    for k, v in options.items():
        env.set_option(k, v)
    if plastic:
        idx = {env.projection(i).type: i for i in range(env.info().n_projections)}
        env.set_plasticity(pkg.STDP_TEACHER, idx[0], idx[11], {pkg.A1_J: 0.0, pkg.A2_JI: -0.01}, 50)
    env.initialize(50, 25)
    info = env.info()
    advance(env, 0, warmup)
    _, ev0 = env.totals()
    t0 = time.perf_counter()
    env.time_steps(steps)
    wall = time.perf_counter() - t0
    _, ev1 = env.totals()
    _, kmean = env.time_steps(100, kernel_mean=True)
    algo = info.algorithmic_bytes_per_step
    out = {"options": dict(options, plasticity="teacher pf->pc" if plastic else None),
           "value": (ev1 - ev0) / wall, "ms_per_step": wall * 1e3 / steps,
           "real_time_factor": steps * 1e-3 / wall, "kernel_mean_ms": kmean, "algorithmic_bytes_per_step": algo,
           "achieved_GBs": algo / (kmean * 1e-3) / 1e9, "synapses_16bit": info.synapses_16bit}
    env.finalize()
    return out


def scaled(pkg, preset, steps, precisions):
    """A tiled scaffold (preset 2: 16x in area, 1.55M cells, 50.7M synapses;
    preset 4: 100x, 9.67M cells, 326M synapses) with the given multiplier storage
    (an entry is a precision, or (precision, index bits)): step time, real-time
    factor, synaptic events/s and the step kernel's algorithmic bandwidth."""
    out = {}
    names = {pkg.WEIGHTS_F32: "fp32", pkg.WEIGHTS_BF16: "bf16", pkg.WEIGHTS_FP8: "fp8"}
    for entry in precisions:
        prec, bits = entry if isinstance(entry, tuple) else (entry, 32)
        name = names[prec] + ("_idx16" if bits == 16 else "")
        env = make_env(pkg, preset)
        env.set_option("weight_precision", prec)
        env.set_option("index_bits", bits)
        env.initialize(50, 25)
        info = env.info()
        advance(env, 0, 50)
        _, ev0 = env.totals()
        t0 = time.perf_counter()
        env.time_steps(steps)
        wall = time.perf_counter() - t0
        _, ev1 = env.totals()
        _, kmean = env.time_steps(50, kernel_mean=True)
        algo = info.algorithmic_bytes_per_step
        out[name] = {"ms_per_step": wall * 1e3 / steps, "real_time_factor": steps * 1e-3 / wall,
                     "synaptic_events_per_s": (ev1 - ev0) / wall, "kernel_mean_ms": kmean,
                     "algorithmic_bytes_per_step": algo, "achieved_GBs": algo / (kmean * 1e-3) / 1e9,
                     "frac": algo / (kmean * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "synapses_16bit": info.synapses_16bit}
        out["neurons"], out["synapses"] = info.total_neurons, info.total_synapses
        env.finalize()
    out["steps"] = steps
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-plastic", action="store_true")
    ap.add_argument("--no-16x", action="store_true")
    ap.add_argument("--steps-16x", type=int, default=200)
    ap.add_argument("--no-100x", action="store_true")
    ap.add_argument("--steps-100x", type=int, default=100)
    ap.add_argument("--batch-instances", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--kernel-steps", type=int, default=100)
    ap.add_argument("--only-scaled", type=int, default=0, choices=(0, 2, 4),
                    help="only run the tiled scaffold of this preset (2: 16x, 4: 100x) and exit")
    ap.add_argument("--scaled-precision", default="fp32", choices=("fp32", "bf16", "fp8"),
                    help="multipliers of the --only-scaled run")
    ap.add_argument("--scaled-index-bits", type=int, default=32, choices=(32, 16),
                    help="column index width of the --only-scaled run")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run the calibration copy kernel on this many MiB and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent instances shard whole per GPU; run one process per GPU)", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as ge
    pkg = ge.load_package()
    if args.only_scaled:
        prec = {"fp32": pkg.WEIGHTS_F32, "bf16": pkg.WEIGHTS_BF16, "fp8": pkg.WEIGHTS_FP8}[args.scaled_precision]
        print(json.dumps(scaled(pkg, args.only_scaled, args.steps, ((prec, args.scaled_index_bits),))))
        return
    if args.copy_probe_mb:
        ms = pkg.copy_probe(args.copy_probe_mb << 20, 20)
        print(json.dumps({"copy_probe_bytes": args.copy_probe_mb << 20, "ms": ms}))
        return

    t0 = time.perf_counter()
    env = make_env(pkg, 1)
    env.set_option("graph", 1)
    env.set_option("graph_steps", 25)
    env.initialize(50, 25)
    info = env.info()
    log(f"built scaffold: {info.total_neurons} cells, {info.total_synapses} synapses, "
        f"{info.n_tasks} wavefront tasks, {time.perf_counter() - t0:.1f} s")

    advance(env, 0, args.warmup)
    # timed region: exactly `steps` exchange steps, synchronized before and after
    sp0, ev0 = env.totals()
    t0 = time.perf_counter()
    gpu_ms = env.time_steps(args.steps)
    wall = time.perf_counter() - t0
    sp1, ev1 = env.totals()
    events = ev1 - ev0
    value = events / wall
    rtf = args.steps * 1e-3 / wall
    # mean duration of the step kernel over runs of back-to-back launches
    _, kmean = env.time_steps(args.kernel_steps, kernel_mean=True)
    algo = info.algorithmic_bytes_per_step
    achieved = algo / (kmean * 1e-3) / 1e9
    # HBM-side bytes per k_step launch from the FETCH_SIZE / WRITE_SIZE counter
    # passes, calibrated against a copy kernel of known size
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_step_scaffold96k_fp32.md")
    if os.path.exists(tpath):
        with open(tpath) as f:
            m = re.search(r"^hbm_bytes_per_launch = (\d+)$", f.read(), re.M)
        traffic = int(m.group(1)) if m else None
    rates = {}
    for i in range(info.n_populations):
        p = env.population(i)
        if p.num:
            rates[p.name.decode()] = round(float(sp1[i] - sp0[i]) / p.num / (args.steps * 1e-3), 3)
    line = {
        "metric": METRIC, "value": value, "unit": "synaptic events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded scaffold generator (reference population sizes, rule-based connectivity)",
        "config": {"workload": "scaffold-96k plain fp32 (1 s = 1000 exchange steps of 10 x 0.1 ms)",
                   "column_index_bits": 32,
                   "neurons": info.total_neurons, "synapses": info.total_synapses,
                   "simulated_ms": args.steps, "input": "4 Hz mossy Poisson + 100 Hz burst to 50 fibres, 300-350 ms"},
        "real_time_factor": rtf,
        "synapses_scanned_per_s": info.total_synapses * args.steps / wall,
        "gpu_ms_per_step": gpu_ms / args.steps,
        "events_in_timed_region": int(events),
        "firing_rates_hz": rates,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_step", "kernel_mean_ms": kmean, "algorithmic_bytes_per_launch": algo},
    }
    env_for_cpu = env
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(env_for_cpu, args.cpu_steps, args.cpu_threads)
    env.finalize()
    if not args.no_batched:  # the same network with 16-bit column offsets (bit-identical results)
        line["plain_fp32_idx16"] = plain_variant(pkg, args.steps, args.warmup, index_bits=16)
    if not args.no_plastic:  # the same network with the reference's Purkinje plasticity rule switched on
        line["plain_fp32_plastic"] = plain_variant(pkg, args.steps, args.warmup, plastic=True)
    if not args.no_batched:
        b = batched(pkg, args.batch_instances, args.batch_steps, 50, threads=min(16, os.cpu_count() or 4))
        b["vs_plain"] = b["value"] / value
        b["vs_plain_synapses_scanned"] = b["synapses_scanned_per_s"] / line["synapses_scanned_per_s"]
        line["batched_bf16"] = b
    if not args.no_16x:
        line["scaffold16x"] = scaled(pkg, 2, args.steps_16x,
                                     (pkg.WEIGHTS_F32, pkg.WEIGHTS_BF16, pkg.WEIGHTS_FP8, (pkg.WEIGHTS_F32, 16),
                                      (pkg.WEIGHTS_BF16, 16), (pkg.WEIGHTS_FP8, 16)))
    if not args.no_100x:
        line["scaffold100x"] = scaled(pkg, 4, args.steps_100x,
                                      ((pkg.WEIGHTS_FP8, 16), pkg.WEIGHTS_FP8, (pkg.WEIGHTS_F32, 16), pkg.WEIGHTS_F32))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
