#!/usr/bin/env python3
"""Benchmark of the exchange-step hot path on one MI355X.

Headline workload: the reference-size cerebellar scaffold from the seeded
scaffold generator with the reference connectome's fan-ins -- 96,737 cells
(88,158 granule, 7,073 glomeruli fed by 2,400 mossy fibres, 69 Purkinje, 219
Golgi, 603 stellate, 603 basket, 12 DCN) and 4.23M synapses, FP32 state and FP32
synaptic multipliers, 4 Hz mossy background plus a 100 Hz burst to a 50-fibre
patch at 300-350 ms.

A bench step is one iteration of the reference's driver loop (main.cu:50-68):
SetInputFreq, then loop_n_steps over T_cpu_sim = 25 exchange steps (25 ms
simulated, each 10 Euler sub-steps of 0.1 ms) with the output spike train handed
over at its end -- the drop-in path a caller of the reference runs.

Prints one JSON line: synaptic events/s (value) and real-time factor of the timed
iterations, the roofline figure of the step kernel (its duration measured with
HIP events recorded on the kernels' stream around each call's step kernels during
the timed region), the CPU baseline (the single-threaded C oracle on a bounded sample
of the same network), and further legs: the device loop alone, the rule-text
density, 16-bit columns, the reference's Purkinje plasticity rule, 64 batched
uneven instances against sequential runs of the same instances, and the 16x /
100x tiled scaffolds.
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
WINDOW = 25          # T_cpu_sim of the reference's main.cu
TIMER_EVERY = 4      # the step-kernel timer samples one driver-loop call in 4
T_PRINT = 50


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_env(pkg, preset, plasticities=0, **over):
    env = pkg.NetworkEnvironment(1, 7, 16, 2, plasticities)
    cfg = pkg.synth_preset(preset)
    for k, v in over.items():
        setattr(cfg, k, v)
    env.describe_synthetic(cfg)
    return env


def proj_nnz(env):
    """Synapses per projection, by the reference's connection type names."""
    names = ["pf_pc", "pf_goc", "pf_sc", "pf_bc", "aa_goc", "aa_pc", "goc_goc", "sc_sc", "bc_bc", "glom_dcn",
             "pc_dcn", "bc_pc", "sc_pc", "glom_grc", "glom_goc", "goc_grc"]
    out = {}
    for i in range(env.info().n_projections):
        p = env.projection(i)
        if p.pre >= 0:
            out[names[p.type] if 0 <= p.type < len(names) else str(p.type)] = int(p.nnz)
    return out


def drop_in_loop(env, t0, iterations, freq):
    """`iterations` driver-loop iterations from step t0: SetInputFreq, then
    loop_n_steps over one window with the output train handed over."""
    t = t0
    for _ in range(iterations):
        env.set_input_freq(freq)
        env.loop_n_steps(t, WINDOW, 0, want_output=True)
        t += WINDOW
    return t


def timed_drop_in(env, t, iterations, freq):
    """Runs `iterations` driver-loop iterations and measures them: wall time (the
    calls return when their steps are done), synaptic events, per-population
    spikes, and the step-kernel time from the HIP events around the step kernels
    of one call in TIMER_EVERY (two event markers in the stream cost the call ~7 us)."""
    info = env.info()
    env.step_timer(TIMER_EVERY)
    sp0, ev0 = env.totals()
    w0 = time.perf_counter()
    t = drop_in_loop(env, t, iterations, freq)
    wall = time.perf_counter() - w0
    kms, launches = env.step_timer(0)
    sp1, ev1 = env.totals()
    ex_steps = iterations * WINDOW
    kmean = kms / max(1, launches)
    algo = info.algorithmic_bytes_per_step
    rates = {}
    for i in range(info.n_populations):
        p = env.population(i)
        if p.num:
            rates[p.name.decode()] = round(float(sp1[i] - sp0[i]) / p.num / (ex_steps * 1e-3), 3)
    return t, {"wall_s": wall, "events": int(ev1 - ev0), "value": (ev1 - ev0) / wall,
               "ms_per_iteration": wall * 1e3 / iterations, "exchange_steps": ex_steps,
               "real_time_factor": ex_steps * 1e-3 / wall,
               "synapses_scanned_per_s": info.total_synapses * ex_steps / wall,
               "kernel_mean_ms": kmean, "kernel_launches_timed": int(launches),
               "algorithmic_bytes_per_step": algo, "achieved_GBs": algo / (kmean * 1e-3) / 1e9 if kmean else None,
               "firing_rates_hz": rates}


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
    log(f"cpu baseline: {steps} steps on 1 thread in {t1:.1f} s")
    evn, tn = timed(threads)
    assert ev1 == evn
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"host_cpu": model, "host_cpus_visible": os.cpu_count(), "value": ev1 / t1, "unit": "synaptic events/s",
            "cores": 1, "kind": "port",
            "sample": f"first {steps} exchange steps ({steps} ms simulated) of the same 96.7k-cell network, "
                      f"{t1:.1f} s on one thread",
            "real_time_factor": steps * 1e-3 / t1, "ms_per_exchange_step": t1 * 1e3 / steps,
            "all_cores": {"value": evn / tn, "unit": "synaptic events/s", "cores": threads,
                          "real_time_factor": steps * 1e-3 / tn, "ms_per_exchange_step": tn * 1e3 / steps,
                          "sample": f"same {steps} steps, OpenMP over {threads} threads, {tn:.1f} s"}}


def variant(pkg, iterations, warmup, preset=1, plastic=False, connectome=None, **options):
    """The plain 96.7k scaffold through the drop-in loop with other options:
    16-bit column indices (same results bit for bit), the rule-text connectome
    (connectome 0: 3.03M synapses), or the reference's parallel-fibre-to-Purkinje
    Teacher rule (A1_J 0, A2_JI -0.01, window 50; SynapticPlasticity.cu:10-16),
    basket-to-Purkinje standing in for the climbing fibre as the teacher."""
    over = {} if connectome is None else {"connectome": connectome}
    env = make_env(pkg, preset, plasticities=1 if plastic else 0, **over)
    for k, v in options.items():
        env.set_option(k, v)
    if plastic:
        idx = {env.projection(i).type: i for i in range(env.info().n_projections)}
        env.set_plasticity(pkg.STDP_TEACHER, idx[0], idx[11], {pkg.A1_J: 0.0, pkg.A2_JI: -0.01}, 50)
    env.initialize(T_PRINT, WINDOW)
    info = env.info()
    freq = [4.0, 4.0]
    t = drop_in_loop(env, 0, warmup, freq)
    _, r = timed_drop_in(env, t, iterations, freq)
    r.pop("firing_rates_hz")
    r.update({"options": dict(options, plasticity="teacher pf->pc" if plastic else None,
                              connectome=connectome if connectome is not None else 1),
              "neurons": int(info.total_neurons), "synapses": int(info.total_synapses),
              "synapses_16bit": int(info.synapses_16bit)})
    env.finalize()
    return r


def batched(pkg, n_inst, steps, warmup, threads, layout=None, profile_steps=0):
    """The batched form: 64 independent instances of uneven size (0.25x..4x of the 96.7k
    scaffold), each with its own seeds and mossy rate U(1, 60) Hz, BF16
    multipliers, advanced together by one step kernel per exchange step
    (hipGraph replays), against sequential plain runs of the same instances: a
    second, identical copy of every instance (its own default layout, which gives
    the same results bit for bit) advanced alone, one after another.  Both copies
    run the same steps from initialization, so their spike totals must agree
    exactly."""
    scales = ([0.25] * 16 + [0.5] * 16 + [1.0] * 16 + [2.0] * 8 + [4.0] * 8)[:n_inst]
    rng = np.random.default_rng(1234)
    rates = rng.uniform(1.0, 60.0, len(scales))
    # autotune the layout on one instance of the largest scale (4x; the 4x
    # instances hold 42 % of the batch's synapses, and at that size the step is
    # throughput-bound like the batch) and use it for all
    # (a given layout skips the tuning: profiles need the same kernels every run)
    if layout is None:
        probe = make_env(pkg, 1, scale=max(scales), burst_n_mf=max(1, int(round(50 * max(scales)))),
                         mf_rate_hz=float(np.median(rates)))
        probe.set_option("weight_precision", pkg.WEIGHTS_BF16)
        tuned = probe.autotune(100)
        probe.finalize()
        log(f"batch: autotuned {tuned}")
        layout = {k: tuned[k] for k in ("lane_max_fanin", "lane_synapses", "graph_steps", "index_bits", "stage_spikes",
                                        "word_groups")}
    else:
        tuned = {"given": dict(layout)}
    envs = [[None] * len(scales), [None] * len(scales)]
    errs = []

    def work(items):
        try:
            for c, i in items:
                e = make_env(pkg, 1, scale=scales[i], seed=1000 + i, stim_seed=5000 + i, mf_rate_hz=float(rates[i]),
                             burst_n_mf=max(1, int(round(50 * scales[i]))))
                e.set_option("weight_precision", pkg.WEIGHTS_BF16)
                e.set_option("neuron_seed", 77 + i)
                if c == 0:  # the batch's copy takes the tuned layout; the sequential plain
                    for k, v in layout.items():  # runs keep each instance's own defaults
                        e.set_option(k, v)
                e.initialize(T_PRINT, WINDOW)
                envs[c][i] = e
        except Exception as ex:  # surfaced below
            errs.append(ex)

    # profile_steps: the batch alone (no sequential copies), advanced one launch
    # per call, without graph capture -- for counter passes
    copies = (0,) if profile_steps else (0, 1)
    order = sorted(((c, i) for c in copies for i in range(len(scales))), key=lambda x: -scales[x[1]])
    parts = [order[k::threads] for k in range(threads)]
    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(p,)) for p in parts]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    if errs:
        raise errs[0]
    log(f"batch: built {len(copies)} x {len(scales)} instances in {time.perf_counter() - t0:.1f} s")
    batch_envs, seq_envs = envs
    if profile_steps:
        b = pkg.Batch(batch_envs)
        w0 = time.perf_counter()
        for _ in range(profile_steps):
            b.run(1)
        wall = time.perf_counter() - w0
        b.destroy()
        for e in batch_envs:
            e.finalize()
        return {"profile_steps": profile_steps, "instances": len(scales), "layout": tuned, "wall_s": wall}
    # sequential plain runs: each instance alone, one after another
    seq_wall = 0.0
    for e in seq_envs:
        e.time_steps(warmup)
    for e in seq_envs:
        w0 = time.perf_counter()
        e.time_steps(steps)
        seq_wall += time.perf_counter() - w0
    log(f"batch: sequential runs {seq_wall * 1e3 / steps:.3f} ms per step")
    b = pkg.Batch(batch_envs)  # replays graphs of the first instance's graph_steps
    b.run(warmup)
    before = sum(e.totals()[1] for e in batch_envs)
    w0 = time.perf_counter()
    ms = b.time(steps)
    wall = time.perf_counter() - w0
    after = sum(e.totals()[1] for e in batch_envs)
    same = all(np.array_equal(x.totals()[0], y.totals()[0]) and x.totals()[1] == y.totals()[1]
               for x, y in zip(batch_envs, seq_envs))
    syn = sum(e.info().total_synapses for e in batch_envs)
    neurons = sum(e.info().total_neurons for e in batch_envs)
    b.destroy()
    for e in batch_envs + seq_envs:
        e.finalize()
    value = (after - before) / wall
    # the sequential copies' events of their timed steps equal the batch's (same
    # instances, same steps), so the ratio of events/s is the ratio of wall times
    seq_value = (after - before) / seq_wall
    return {"value": value, "unit": "synaptic events/s", "instances": len(scales),
            "scales": {str(s): scales.count(s) for s in sorted(set(scales))},
            "synapses_scanned_per_s": syn * steps / wall,
            "neurons": int(neurons), "synapses": int(syn), "steps": steps, "warmup": warmup,
            "ms_per_step": wall * 1e3 / steps, "gpu_ms_per_step": ms / steps, "dtype": "bf16 multipliers, f32 state",
            "aggregate_real_time_factor": len(scales) * steps * 1e-3 / wall, "autotuned": tuned,
            "sequential": {"value": seq_value, "ms_per_step": seq_wall * 1e3 / steps,
                           "synapses_scanned_per_s": syn * steps / seq_wall,
                           "spike_totals_identical_to_batch": bool(same)},
            "vs_sequential": seq_wall / wall}


def scaled(pkg, preset, steps, precisions):
    """A tiled scaffold (preset 2: 16x in area, 1.55M cells, 67M synapses;
    preset 4: 100x, 9.67M cells, 421M synapses) with the given multiplier storage
    (an entry is a precision, or (precision, index bits)): step time, real-time
    factor, synaptic events/s and the step kernel's algorithmic bandwidth."""
    out = {}
    names = {pkg.WEIGHTS_F32: "fp32", pkg.WEIGHTS_BF16: "bf16", pkg.WEIGHTS_FP8: "fp8"}
    for entry in precisions:
        prec, bits = entry if isinstance(entry, tuple) else (entry, 32)
        name = names[prec] + ("_idx16" if bits == 16 else "")
        t0 = time.perf_counter()
        env = make_env(pkg, preset)
        env.set_option("weight_precision", prec)
        env.set_option("index_bits", bits)
        env.initialize(T_PRINT, WINDOW)
        info = env.info()
        log(f"scaffold preset {preset} {name}: built in {time.perf_counter() - t0:.1f} s")
        env.time_steps(50)
        _, ev0 = env.totals()
        env.step_timer(1)
        w0 = time.perf_counter()
        env.time_steps(steps)
        wall = time.perf_counter() - w0
        kms, launches = env.step_timer(0)
        _, ev1 = env.totals()
        kmean = kms / max(1, launches)
        algo = info.algorithmic_bytes_per_step
        out[name] = {"ms_per_step": wall * 1e3 / steps, "real_time_factor": steps * 1e-3 / wall,
                     "synaptic_events_per_s": (ev1 - ev0) / wall, "kernel_mean_ms": kmean,
                     "algorithmic_bytes_per_step": algo, "achieved_GBs": algo / (kmean * 1e-3) / 1e9,
                     "frac": algo / (kmean * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "layout_bytes_per_step": info.stream_bytes_per_step,
                     "layout_frac": info.stream_bytes_per_step / (kmean * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "synapses_16bit": info.synapses_16bit, "word_group_synapses": info.word_group_synapses}
        out["neurons"], out["synapses"] = info.total_neurons, info.total_synapses
        env.finalize()
    out["steps"] = steps
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40,
                    help="timed driver-loop iterations of 25 exchange steps each (40 = 1 s simulated)")
    ap.add_argument("--warmup", type=int, default=4, help="untimed driver-loop iterations before them")
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-scaled", action="store_true")
    ap.add_argument("--full", action="store_true",
                    help="tiled scaffolds in every storage form (default: FP8 with 16-bit columns only)")
    ap.add_argument("--steps-16x", type=int, default=200)
    ap.add_argument("--steps-100x", type=int, default=100)
    ap.add_argument("--batch-instances", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--device-steps", type=int, default=1000)
    ap.add_argument("--only-scaled", type=int, default=0, choices=(0, 2, 4),
                    help="only run the tiled scaffold of this preset (2: 16x, 4: 100x) and exit")
    ap.add_argument("--scaled-precision", default="fp32", choices=("fp32", "bf16", "fp8"),
                    help="multipliers of the --only-scaled run")
    ap.add_argument("--scaled-index-bits", type=int, default=32, choices=(32, 16),
                    help="column index width of the --only-scaled run")
    ap.add_argument("--only-batched", action="store_true",
                    help="only run the batched leg (64 instances of 0.25x-4x) and print its result")
    ap.add_argument("--batch-layout", type=json.loads, default=None,
                    help="layout options of the batched leg as JSON (skips the autotuner), e.g. "
                         "'{\"lane_max_fanin\": 128, \"lane_synapses\": 32, \"graph_steps\": 50, "
                         "\"index_bits\": 16, \"stage_spikes\": 1}'")
    ap.add_argument("--batch-profile-steps", type=int, default=0,
                    help="with --only-batched: build the batch alone and advance it this many steps, one launch "
                         "per call and no graph capture (for counter passes)")
    ap.add_argument("--only-variant", default=None, choices=("rule_text", "idx16", "plastic"),
                    help="only run this variant of the plain 96.7k drop-in loop and print its result")
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
    if args.only_batched:
        print(json.dumps(batched(pkg, args.batch_instances, args.batch_steps, 50, layout=args.batch_layout,
                                 profile_steps=args.batch_profile_steps,
                                 threads=min(16, os.cpu_count() or 4))))
        return
    if args.only_variant:
        kw = {"rule_text": {"connectome": 0}, "idx16": {"index_bits": 16}, "plastic": {"plastic": True}}
        print(json.dumps(variant(pkg, args.steps, args.warmup, **kw[args.only_variant])))
        return
    if args.copy_probe_mb:
        ms = pkg.copy_probe(args.copy_probe_mb << 20, 20)
        print(json.dumps({"copy_probe_bytes": args.copy_probe_mb << 20, "ms": ms}))
        return

    t0 = time.perf_counter()
    env = make_env(pkg, 1)
    env.set_option("graph_steps", WINDOW)  # one run (and one run-end launch) per call
    env.initialize(T_PRINT, WINDOW)
    info = env.info()
    log(f"built scaffold: {info.total_neurons} cells, {info.total_synapses} synapses, "
        f"{info.n_tasks} wavefront tasks, {time.perf_counter() - t0:.1f} s")
    freq = [4.0, 4.0]  # CalcFreq's background rate (main.cu:7); the burst patch keeps its own
    t = drop_in_loop(env, 0, args.warmup, freq)
    # timed region: exactly `steps` driver-loop iterations; each call returns when
    # its exchange steps are done and its output train is in host memory
    t, r = timed_drop_in(env, t, args.steps, freq)
    log(f"timed: {args.steps} iterations, {r['ms_per_iteration']:.3f} ms each, kernel mean {r['kernel_mean_ms']:.5f} ms")
    # HBM-side bytes per k_step launch from the FETCH_SIZE / WRITE_SIZE counter
    # passes, calibrated against a copy kernel of known size
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_step_scaffold96k_fp32.md")
    if os.path.exists(tpath):
        with open(tpath) as f:
            m = re.search(r"^hbm_bytes_per_launch = (\d+)$", f.read(), re.M)
        traffic = int(m.group(1)) if m else None
    kmean = r["kernel_mean_ms"]
    achieved = r["achieved_GBs"]
    line = {
        "metric": METRIC, "value": r["value"], "unit": "synaptic events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_iteration"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded scaffold generator (reference population sizes, reference connectome fan-ins)",
        "config": {"workload": "scaffold-96k plain fp32, drop-in driver loop",
                   "step": f"one loop_n_steps call of T_cpu_sim = {WINDOW} exchange steps ({WINDOW} ms simulated, "
                           f"10 x 0.1 ms Euler sub-steps each) with SetInputFreq before it and the output spike "
                           f"train handed over after it",
                   "exchange_steps_timed": r["exchange_steps"], "simulated_ms": r["exchange_steps"],
                   "first_timed_exchange_step": args.warmup * WINDOW,
                   "column_index_bits": 32, "neurons": info.total_neurons, "synapses": info.total_synapses,
                   "word_groups": info.word_groups, "word_group_synapses": info.word_group_synapses,
                   "synapses_per_projection": proj_nnz(env),
                   "input": "4 Hz mossy Poisson + 100 Hz burst to 50 fibres, 300-350 ms",
                   "launches": "per call: 25 direct step-kernel launches, then one run-end launch (statistics, "
                               "output train, step counters)"},
        "real_time_factor": r["real_time_factor"],
        "synapses_scanned_per_s": r["synapses_scanned_per_s"],
        "events_in_timed_region": r["events"],
        "firing_rates_hz": r["firing_rates_hz"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "kernel": "k_step", "kernel_mean_ms": kmean, "launches_timed": r["kernel_launches_timed"],
                     "algorithmic_bytes_per_launch": r["algorithmic_bytes_per_step"],
                     # the bytes the stored layout streams per launch (word groups read one
                     # 16-byte record per presynaptic spike word instead of an index and a
                     # multiplier per synapse), and the fraction of the peak they reach
                     "layout_bytes_per_launch": info.stream_bytes_per_step,
                     "layout_frac": info.stream_bytes_per_step / (kmean * 1e-3) / 1e9 / HBM_PEAK_GBS if kmean else None,
                     "timing": f"HIP events on the step kernels' stream around the 25 step kernels of one call in "
                               f"{TIMER_EVERY} (the first included), timed region"},
    }
    # the device loop alone (runs of 25 steps, no output train, no host calls between)
    env.step_timer(1)
    w0 = time.perf_counter()
    env.time_steps(args.device_steps)
    wall = time.perf_counter() - w0
    kms, launches = env.step_timer(0)
    line["device_loop"] = {"exchange_steps": args.device_steps, "ms_per_exchange_step": wall * 1e3 / args.device_steps,
                           "real_time_factor": args.device_steps * 1e-3 / wall,
                           "kernel_mean_ms": kms / max(1, launches)}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(env, args.cpu_steps, args.cpu_threads)
    env.finalize()
    if not args.no_variants:
        line["plain_fp32_rule_text_density"] = variant(pkg, args.steps, args.warmup, connectome=0)
        log("variant: rule-text density done")
        line["plain_fp32_idx16"] = variant(pkg, args.steps, args.warmup, index_bits=16)
        log("variant: 16-bit columns done")
        line["plain_fp32_plastic"] = variant(pkg, args.steps, args.warmup, plastic=True)
        log("variant: plasticity done")
    if not args.no_batched:
        b = batched(pkg, args.batch_instances, args.batch_steps, 50, threads=min(16, os.cpu_count() or 4))
        b["vs_plain"] = b["value"] / line["value"]
        b["vs_plain_synapses_scanned"] = b["synapses_scanned_per_s"] / line["synapses_scanned_per_s"]
        line["batched_bf16"] = b
        log("batched done")
    if not args.no_scaled:
        full = args.full
        line["scaffold16x"] = scaled(pkg, 2, args.steps_16x,
                                     (pkg.WEIGHTS_F32, pkg.WEIGHTS_BF16, pkg.WEIGHTS_FP8, (pkg.WEIGHTS_F32, 16),
                                      (pkg.WEIGHTS_BF16, 16), (pkg.WEIGHTS_FP8, 16)) if full
                                     else ((pkg.WEIGHTS_FP8, 16),))
        line["scaffold100x"] = scaled(pkg, 4, args.steps_100x,
                                      ((pkg.WEIGHTS_FP8, 16), pkg.WEIGHTS_FP8, (pkg.WEIGHTS_F32, 16), pkg.WEIGHTS_F32)
                                      if full else ((pkg.WEIGHTS_FP8, 16),))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
