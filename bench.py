#!/usr/bin/env python3
"""Benchmark of the HVC growth-trial hot path on one MI355X.

A step is one growth trial (500 ms, 24 999 neuron steps, STDP on) of the
20 000 HVC-RA + 5 500 HVC-I network (configs.c3): 100 RA->RA candidates per
RA by distance, delays U(0, 10) ms, RK4 form (HH2_buffer::Runge4_step for
HVC-RA; HVC-I use HHI_buffer::DP8_step because HHI_buffer::Runge4_step
diverges at the reference's dt = 0.02 ms), in the precision form of that
configuration: FP32 neuron state, FP64 clock, spike times and STDP.

    python bench.py [--steps K] [--warmup W] [--precision fp32|fp64] [--config c1|c2|c3]
    python bench.py --config c4 [--precision fp32|fp64]     # batched form
    python bench.py --config c5 [--precision fp32|fp64]     # served form

Prints one JSON line: neuron-updates/s = (N_RA + N_I) * 24 999 * K / wall time,
with inputs resident on the device, plus the roofline of the dominant kernel
(HIP-event durations on the engine stream), a CPU baseline (the oracle's
restatement of the same trial, one host thread, on a bounded sample) and, for
FP32, the per-trial spike-time error against the FP64 form (bit-exact with the
oracle) and the FP64 form's own throughput on the same trials.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# the FP32 vector rate of unpacked instructions (one FMA per lane), the ceiling of
# kernels built without packed FP32 code generation, as the engine's are
FP32_UNPACKED_TFLOPS = 78.6


def unpacked(achieved, precision):
    """FP32: the achieved rate against the unpacked ceiling (None for FP64)."""
    return None if precision == 0 else {"ceiling": FP32_UNPACKED_TFLOPS, "frac": achieved / FP32_UNPACKED_TFLOPS}

# HBM bytes per launch of the integration kernel, from the rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes (scaled by an elementwise kernel of known traffic) recorded in
# profiles/integrate_c3_hbm_traffic.txt; keyed by (config, precision, HVC-I layout)
TRAFFIC_PROFILED = {("c3", 1, 3): 7.244e6, ("c3", 0, 3): 7.680e6}
# HBM bytes per launch of the delivery kernel (same passes), keyed by (config, delivery form: 0 sweep, 1 lists)
DELIVER_TRAFFIC_PROFILED = {("c3", 0): 1.278e6, ("c3", 1): 1.873e6}
DELIVER_TRAFFIC_NOTE = "profiles/integrate_c3_hbm_traffic.txt"
HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--ra-model", type=int, default=1, help="0 Debraband (drop-in), 1 RK4 (HH2_buffer)")
    ap.add_argument("--i-model", type=int, default=0, help="0 DP8, 1 RK4")
    ap.add_argument("--cpu-sample-ms", type=float, default=50.0,
                    help="trial length of the CPU-baseline sample (same network)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--dropin-trials", type=int, default=2,
                    help="timed trials of the drop-in pairing (Debraband + DP8) in FP32 and FP64; 0 skips")
    ap.add_argument("--nets", type=int, default=64, help="c4: number of networks")
    ap.add_argument("--check-nets", type=int, default=4, help="c4: networks re-run alone for comparison")
    ap.add_argument("--trial-ms", type=float, default=None, help="trial length (default: the config's)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="read and write this many MiB in one elementwise kernel before the trials "
                         "(calibrates the HBM byte counters)")
    ap.add_argument("--tune-windows", type=int, default=400, help="coupling windows per autotuner candidate (0: no tuning)")
    ap.add_argument("--shape", default=None,
                    help="launch shape integrate_block,i_lanes,deliver_block,sync_block[,ra_lanes[,deliver_lists]] "
                         "(skips the autotuner)")
    ap.add_argument("--driver-trials", type=int, default=2,
                    help="c3: growth-loop trials through the C++ driver (bookkeeping included); 0 skips")
    ap.add_argument("--requests", type=int, default=16, help="c5: requests of the first load point")
    ap.add_argument("--trials", type=int, default=50, help="c5: chain-test trials per request")
    ap.add_argument("--rate", type=float, default=20.0, help="c5: arrivals per second of the first load point")
    ap.add_argument("--sustain-requests", type=int, default=100,
                    help="c5: requests of the second load point (0 skips it)")
    ap.add_argument("--sustain-rate", type=float, default=1.0,
                    help="c5: arrivals per second of the second load point (one the device sustains)")
    ap.add_argument("--serve", default="async", choices=["async", "rounds"],
                    help="c5: hvc_submit / hvc_poll (async) or rounds driven by the caller (hvc_pool_run)")
    ap.add_argument("--check-requests", type=int, default=3, help="c5: requests re-run alone in FP64")
    ap.add_argument("--check-trials", type=int, default=10, help="c5: trials per checked request")
    ap.add_argument("--variants", type=int, default=2, help="c5: networks built per size")
    ap.add_argument("--hbm-gib", type=float, default=8.0, help="c5: device memory of the serving pool")
    ap.add_argument("--c2-trials", type=int, default=100, help="c2: growth trials through the driver")
    ap.add_argument("--c2-fp64", type=int, default=1, help="c2: 1 runs the FP64 form (the config's), 0 follows "
                                                          "--precision")
    ap.add_argument("--activity-window", type=int, default=25,
                    help="c2: trials of the activity check (and replacement cadence)")
    return ap.parse_args()


def cpu_baseline(net, kw, args):
    """The oracle (C restatement of the trial, -O2) on the host cores: the neuron
    steps of each time step on OpenMP threads (every core the process may use),
    deliveries and STDP in the reference's order on one thread; the same
    network, a shortened growth trial whose training kick falls at a fifth of
    it (as 100 ms does in a 500 ms trial), after an untimed first trial."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleSim, oracle
    sample_ms = args.cpu_sample_ms
    sim = OracleSim(net, ra_model=args.ra_model, i_model=args.i_model, trial_ms=sample_ms,
                    seed=kw.get("seed", (1, 0)))
    steps = int(round(sample_ms / 0.02)) - 1
    sim.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=0.2 * sample_ms)
    t = time.perf_counter()
    sim.run_trial(mode=0, training=True, trial_index=1, first_trial=False, mean_injection_time=0.2 * sample_ms)
    el = time.perf_counter() - t
    sim.close()
    threads = oracle().ora_threads()
    return {"value": (net.n_ra + net.n_i) * steps / el, "unit": "neuron-updates/s", "cores": threads,
            "kind": "port",
            "sample": "oracle growth trial of %g ms (%d steps, kick at %g ms) on the same %d RA + %d I network, "
                      "%.2f s wall; neuron steps on %d OpenMP threads (%d host CPUs visible), delivery and STDP "
                      "on one thread, -O2" % (sample_ms, steps, 0.2 * sample_ms, net.n_ra, net.n_i, el, threads,
                                              os.cpu_count() or 0)}


def bench_driver(net, kw, args, precision, tuning):
    """The growth loop of chain_growth through the C++ driver: trials with
    random injection times plus the bookkeeping after each (activity history,
    maturation, the Erest / GCa patch; no replacement happens this early)."""
    from unsupervised_growth_amd import configs
    from unsupervised_growth_amd.driver import GrowthDriver
    from unsupervised_growth_amd.engine import make_params
    from unsupervised_growth_amd.network import MaturationParams
    ra, ii, mid, top = configs.c3_geometry()
    params = make_params(ra_model=args.ra_model, i_model=args.i_model, precision=precision, **kw)
    with GrowthDriver(net, ra, ii, mid, top, MaturationParams(), params=params) as drv:
        drv.set_tuning(**tuning)
        drv.trial()                    # warm-up (first trial also matures the mature-valued neurons)
        wall, reps = 0.0, []
        for _ in range(args.driver_trials):
            t0 = time.perf_counter()
            reps.append(drv.trial())
            wall += time.perf_counter() - t0
    steps = int(round(kw["trial_ms"] / 0.02)) - 1
    return {"value": (net.n_ra + net.n_i) * steps * args.driver_trials / wall, "unit": "neuron-updates/s",
            "ms_per_step": 1000.0 * wall / args.driver_trials, "trials": args.driver_trials,
            "spikes_per_trial": sum(r.soma_spikes + r.i_spikes for r in reps) / len(reps),
            "note": "growth loop through the C++ driver (hvc_driver_trial): trial with a random injection time, "
                    "spike read-back, check_neuron_activity, maturation update and the Erest/GCa patch"}


def bench_growth_c2(args):
    """The HVC-scale growth run (configs.c2, FP64 by default): the dense 2000 RA
    + 550 I network grown trial after trial through the C++ driver (random
    injection times, check_neuron_activity over a window of 25 trials,
    maturation, replacement of silent neurons with resampled connections, the
    Erest / GCa patch), every step of the loop inside the timed region."""
    from unsupervised_growth_amd import configs
    from unsupervised_growth_amd.driver import GrowthDriver
    from unsupervised_growth_amd.engine import make_params
    from unsupervised_growth_amd.network import MaturationParams
    t_build = time.perf_counter()
    net, kw = configs.c2()
    ra, ii, mid, top = configs.c2_geometry()
    t_build = time.perf_counter() - t_build
    if args.trial_ms:
        kw["trial_ms"] = args.trial_ms
    precision = 0 if args.c2_fp64 else (0 if args.precision == "fp64" else 1)
    params = make_params(ra_model=args.ra_model, i_model=args.i_model, precision=precision, **kw)
    n = args.c2_trials
    with GrowthDriver(net, ra, ii, mid, top, MaturationParams(), params=params,
                      activity_window=args.activity_window) as drv:
        tuning = None
        if args.shape:
            v = [int(x) for x in args.shape.split(",")]
            drv.set_tuning(v[0], v[1], v[2], v[3], v[4] if len(v) > 4 else 1, v[5] if len(v) > 5 else 1)
        elif args.tune_windows > 0:
            tuning = drv.autotune(windows=args.tune_windows)   # before the timed region
        reps = []
        t0 = time.perf_counter()
        for k in range(n):
            reps.append(drv.trial())
            if (k + 1) % 10 == 0:
                print("c2: %d trials, %.1f s" % (k + 1, time.perf_counter() - t0), file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
    steps = int(round(kw["trial_ms"] / 0.02)) - 1
    replaced = [(k, r.replaced) for k, r in enumerate(reps) if r.replaced]
    line = {
        "metric": "RK4 neuron-updates/s, HVC-scale growth through the driver (replacement included)",
        "value": (net.n_ra + net.n_i) * steps * n / wall, "unit": "neuron-updates/s", "n_gpus": 1,
        "steps": n, "warmup": 0, "ms_per_step": 1000.0 * wall / n, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if precision == 0 else "f32",
        "data": "synthetic network (distance rules, seeded), counter-based noise streams",
        "config": {"workload": "growth of %d HVC-RA + %d HVC-I, dense RA->RA candidates (%d), lognormal delays "
                               "(median 2 ms, sigma 0.5, <= 15 ms), %d trials of %g ms, RA %s + I DP8, "
                               "activity window and replacement cadence %d trials"
                               % (net.n_ra, net.n_i, net.n_candidates, n, kw["trial_ms"],
                                  "RK4-3/8" if args.ra_model else "Debraband", args.activity_window),
                   "config": "c2", "network_build_s": round(t_build, 1)},
        "replaced_by_trial": replaced,
        "first_replacement_trial": replaced[0][0] if replaced else None,
        "spikes_per_trial": sum(r.soma_spikes + r.i_spikes for r in reps) / max(len(reps), 1),
        "matured": sum(r.matured for r in reps),
        "tuning": tuning,
    }
    print(json.dumps(line), flush=True)


def deliver_bytes(net, results, windows, avg_ms, key):
    """The delivery kernel against its bytes: algorithmic bytes per launch (every
    edge a spike delivers: 4 B target + 8 B delay + 8 B weight read, 16 B inbox
    record written) from the timed trials' spikes and the network's out-degrees
    (unsaturated RA->RA rows deliver to every candidate, NetworkGrowthSimulator.cpp:9010-9030),
    and the HBM bytes rocprofv3 counted per launch for this workload."""
    import numpy as np
    ra_deg = np.diff(net.ra2i_ptr) + np.diff(net.rr_ptr)
    i_deg = np.diff(net.i2ra_ptr)
    edges = 0
    for r in results:
        edges += int(np.dot(np.asarray([len(x) for x in r.soma]), ra_deg))
        edges += int(np.dot(np.asarray([len(x) for x in r.interneuron]), i_deg))
    alg = 36.0 * edges / (len(results) * windows)
    meas = DELIVER_TRAFFIC_PROFILED.get(key)
    s = avg_ms / 1000.0
    out = {"avg_launch_ms": avg_ms, "delivered_edges_per_trial": edges / len(results),
           "algorithmic_bytes_per_launch": alg, "algorithmic_GBps": alg / s / 1e9 if s > 0 else None,
           "hbm_bytes_per_launch": meas, "hbm_GBps": meas / s / 1e9 if (meas and s > 0) else None,
           "peak_GBps": HBM_PEAK_GBPS}
    if meas:
        out["note"] = ("HBM bytes from rocprofv3 FETCH_SIZE / WRITE_SIZE per k_deliver dispatch, scaled by a kernel "
                       "of known traffic: %s" % DELIVER_TRAFFIC_NOTE)
    return out


def bench_dropin(net, kw, args, tuning):
    """The drop-in pairing of the reference's drivers (HH2_test Debraband for
    HVC-RA, HHI_buffer DP8 for HVC-I; NetworkGrowthSimulator.cpp:8943, 9223) on
    the same network, in FP32 and FP64: growth trials per second."""
    from unsupervised_growth_amd.engine import Engine
    out = {}
    steps = int(round(kw["trial_ms"] / 0.02)) - 1
    for name, prec in (("fp32", 1), ("fp64", 0)):
        with Engine(ra_model=0, i_model=0, precision=prec, **kw) as e:
            e.upload(net)
            e.set_tuning(**dict(tuning, ra_lanes=0))
            e.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=100.0)
            wall = 0.0
            for k in range(args.dropin_trials):
                t0 = time.perf_counter()
                e.run_trial(mode=0, training=True, trial_index=1 + k, mean_injection_time=100.0)
                wall += time.perf_counter() - t0
        out[name] = {"value": (net.n_ra + net.n_i) * steps * args.dropin_trials / wall, "unit": "neuron-updates/s",
                     "ms_per_step": 1000.0 * wall / args.dropin_trials, "trials": args.dropin_trials}
    out["note"] = "HVC-RA HH2_test::Debraband_step + HVC-I HHI_buffer::DP8_step (what the reference's drivers run)"
    return out


def apply_shape(eng, args):
    """The launch shape: given on the command line, else chosen by the autotuner."""
    if args.shape:
        v = [int(x) for x in args.shape.split(",")]
        eng.set_tuning(integrate_block=v[0], i_lanes=v[1], deliver_block=v[2], sync_block=v[3],
                       ra_lanes=v[4] if len(v) > 4 else 1, deliver_lists=v[5] if len(v) > 5 else 1)
    elif args.tune_windows > 0:
        shape, ms = eng.autotune(windows=args.tune_windows)
        eng.tune_ms = ms
    return eng.tuning()


def bench_batch(args):
    """Batched form (configs.c4): all networks advance together in one context."""
    from unsupervised_growth_amd import configs, opcount
    from unsupervised_growth_amd.engine import Engine
    from unsupervised_growth_amd.spikes import compare_results

    t_build = time.perf_counter()
    nets, keys, kw = configs.c4(n_nets=args.nets)
    t_build = time.perf_counter() - t_build
    precision = 0 if args.precision == "fp64" else 1
    eng = Engine(ra_model=args.ra_model, i_model=args.i_model, precision=precision, **kw)
    eng.upload_batch(nets, keys)
    t_tune = time.perf_counter()
    tuning = apply_shape(eng, args)
    tune_ms = getattr(eng, "tune_ms", None)
    t_tune = time.perf_counter() - t_tune
    steps_per_trial = int(round(kw["trial_ms"] / 0.02)) - 1
    n_neurons = sum(n.n_ra + n.n_i for n in nets)

    def trials(tr):
        return [dict(mode=0, training=True, trial_index=tr, first_trial=(tr == 0), mean_injection_time=100.0)
                for _ in nets]

    trial = 0
    for _ in range(args.warmup):
        eng.run_batch(trials(trial))
        trial += 1
    syn0 = [eng.download_synapses(b) for b in range(len(nets))]
    first_timed = trial
    eng.set_timing(True)
    t0 = time.perf_counter()
    results = []
    for _ in range(args.steps):
        results.append(eng.run_batch(trials(trial)))
        trial += 1
    wall = time.perf_counter() - t0
    kern = {name: eng.kernel_timing(name)[0] for name in ("integrate", "deliver", "sync", "decay")}
    eng.close()

    # network by network against the unbatched form (same trial, same synapse state)
    check_nets = list(range(min(args.check_nets, len(nets))))
    worst_alone = {"max_abs_dt_ms": 0.0, "unpaired_spikes": 0, "spikes": 0, "networks": len(check_nets)}
    for b in check_nets:
        with Engine(ra_model=args.ra_model, i_model=args.i_model, precision=precision, seed=keys[b], **kw) as e1:
            e1.upload(nets[b])
            e1.set_tuning(**tuning)          # the batch's launch shape: the FP32 rounding follows the layout
            e1.upload_synapses(syn0[b])
            r1 = e1.run_trial(mode=0, training=True, trial_index=first_timed, first_trial=(first_timed == 0),
                              mean_injection_time=100.0)
        c = compare_results(r1, results[0][b], window=1.0)
        worst_alone["max_abs_dt_ms"] = max(worst_alone["max_abs_dt_ms"], max(v["max_abs_dt_ms"] for v in c.values()))
        worst_alone["unpaired_spikes"] += sum(v["unpaired"] for v in c.values())
        worst_alone["spikes"] += sum(v["spikes_ref"] for v in c.values())
    precision_check = None
    if precision == 1:
        with Engine(ra_model=args.ra_model, i_model=args.i_model, precision=0, **kw) as e64:
            e64.upload_batch(nets, keys)
            for b in range(len(nets)):
                e64.upload_synapses(syn0[b], b)
            r64 = e64.run_batch(trials(first_timed))
        tot = {"max_abs_dt_ms": 0.0, "unpaired_spikes": 0, "paired_within_one_step": 0, "spikes": 0,
               "networks_exceeding_one_step": 0}
        for b in range(len(nets)):
            c = compare_results(r64[b], results[0][b], window=1.0)
            worst = max(v["max_abs_dt_ms"] for v in c.values())
            tot["max_abs_dt_ms"] = max(tot["max_abs_dt_ms"], worst)
            tot["networks_exceeding_one_step"] += int(worst > 0.02 + 1e-9)
            tot["unpaired_spikes"] += sum(v["unpaired"] for v in c.values())
            tot["paired_within_one_step"] += sum(v["paired_within_one_step"] for v in c.values())
            tot["spikes"] += sum(v["spikes_ref"] for v in c.values())
        precision_check = dict(reference="FP64 batched form (bit-exact with the oracle), same trial", dt_ms=0.02, **tot)

    windows = int(round(kw["trial_ms"] / 0.25))
    ra_flops = opcount.flops_per_step("ra", args.ra_model, precision)
    i_flops = opcount.flops_per_step("i", args.i_model, precision)
    flops_launch = sum(ra_flops * n.n_ra + i_flops * n.n_i for n in nets) * steps_per_trial / windows
    peak = FP64_PEAK_TFLOPS if precision == 0 else FP32_PEAK_TFLOPS
    achieved = flops_launch / (kern["integrate"] / 1000.0) / 1e12 if kern["integrate"] > 0 else 0.0
    line = {
        "metric": "RK4 neuron-updates/s, batched form (all networks together)",
        "value": n_neurons * steps_per_trial * args.steps / wall, "unit": "neuron-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == 0 else "f32",
        "data": "synthetic networks (distance rules, seeded), counter-based noise streams",
        "config": {"workload": "batched growth trials: %d networks, %d HVC-RA + %d HVC-I in total, N_RA %d..%d, "
                               "%g ms trials, RA %s + I DP8, STDP on"
                               % (len(nets), sum(n.n_ra for n in nets), sum(n.n_i for n in nets),
                                  min(n.n_ra for n in nets), max(n.n_ra for n in nets), kw["trial_ms"],
                                  "RK4-3/8" if args.ra_model else "Debraband"),
                   "config": "c4", "network_build_s": round(t_build, 1)},
        "roofline": {"bound": "valu", "kernel": "integrate", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "flops_per_launch": flops_launch,
                     "unpacked_fp32": unpacked(achieved, precision),
                     "avg_launch_ms": kern["integrate"],
                     "kernels_ms_per_trial": {k: v * (1 if k in ("decay", "sync") else windows)
                                              for k, v in kern.items()}},
        "batch_vs_alone": worst_alone,
        "autotune": {"tuning": tuning, "s": round(t_tune, 2),
                     "ms_per_window": None if tune_ms is None else [float(x) for x in tune_ms]},
    }
    if precision_check is not None:
        line["precision_check"] = precision_check
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Served form (configs c5): Poisson arrivals of chain-test requests on grown
    networks of mixed size, continuously batched on one pool."""
    import numpy as np
    from unsupervised_growth_amd import configs
    from unsupervised_growth_amd.engine import Pool, network_sizes, pool_caps
    from unsupervised_growth_amd.serve import Request, Server

    t_build = time.perf_counter()
    nets = configs.c5_networks(variants=args.variants)
    t_build = time.perf_counter() - t_build
    all_nets = [n for v in nets.values() for n in v]
    sizes = [network_sizes(n) for n in all_nets]
    indeg = 0
    for n in all_nets:
        indeg = max(indeg, int(np.bincount(n.i2ra_tgt, minlength=n.n_ra).max(initial=0)),
                    int(np.bincount(n.ra2i_tgt, minlength=max(n.n_i, 1)).max(initial=0)),
                    int(np.bincount(n.rr_tgt, minlength=n.n_ra).max(initial=0)))
    # capacities from the HBM budget: neurons and connections in the workload's proportions
    per_ra = {k: max(s[j] / s[0] for s in sizes) * 1.25 for j, k in ((1, "i"), (2, "rr"), (3, "a2i"), (4, "i2ra"))}
    precision = 0 if args.precision == "fp64" else 1
    kw = dict(trial_ms=500.0, precision=precision, ra_model=args.ra_model, i_model=args.i_model)
    probe = pool_caps(1, 1, 1, 1, 1, 1, indeg + 8, 128)
    from unsupervised_growth_amd.engine import make_params
    import ctypes
    from unsupervised_growth_amd import _abi
    lib = _abi.load()
    prm = make_params(**kw)
    unit = ctypes.c_size_t()
    lib.hvc_pool_footprint(ctypes.byref(prm), ctypes.byref(probe), 1000, int(1000 * per_ra["i"]),
                           int(1000 * per_ra["rr"]), int(1000 * per_ra["a2i"]), int(1000 * per_ra["i2ra"]),
                           ctypes.byref(unit), None)
    max_ra = int(args.hbm_gib * (1 << 30) / (unit.value / 1000.0))
    caps = pool_caps(64, max_ra, int(max_ra * per_ra["i"]), int(max_ra * per_ra["rr"]), int(max_ra * per_ra["a2i"]),
                     int(max_ra * per_ra["i2ra"]), indeg + 8, 128)
    from unsupervised_growth_amd.serve import AsyncServer
    from unsupervised_growth_amd.spikes import compare_results

    def serve(rate, n_req, seed):
        reqs0 = configs.c5_requests(nets, n_requests=n_req, rate=rate, trials=args.trials, seed=seed)
        with Pool(caps, **kw) as pool:
            # the pool's launch shape (its networks change: no autotuner); FP32 by
            # default one thread per neuron, the fastest layout once the pool holds
            # more neurons than fill the chip at eight lanes per HVC-I neuron
            if args.shape:
                apply_shape(pool, args)
            elif precision == 1:
                pool.set_tuning(integrate_block=64, i_lanes=0, deliver_block=64, sync_block=256, ra_lanes=0,
                                deliver_lists=1)
            reqs = [Request(rid=k, arrival=a, net=n, key=key, onsets=on, mode=1, first_trial_index=0)
                    for k, (a, n, key, on) in enumerate(reqs0)]
            if args.serve == "async":
                stats = AsyncServer(pool).run(reqs, keep_results=True)
            else:
                stats = Server(pool).run(reqs, realtime=True, keep_results=True)
            stats["pool_gib"] = pool.pool_bytes() / (1 << 30)
            stats["tuning"] = pool.tuning()
        lat = sorted(r.finished - r.arrival for r in reqs)
        pt = {"offered_rate": rate, "requests": n_req, "requests_per_s": n_req / stats["wall_s"],
              "neuron_updates_per_s": stats["neuron_updates_per_s"], "wall_s": stats["wall_s"],
              "rounds": stats["rounds"],
              "latency_s": {"p50": stats["latency_p50"], "p99": stats["latency_p99"], "max": stats["latency_max"],
                            "min": lat[0]}}
        return reqs, pt, stats

    points = []
    reqs_sus = None
    for rate, n_req, seed in ((args.rate, args.requests, 7), (args.sustain_rate, args.sustain_requests, 8)):
        if n_req <= 0:
            continue
        reqs, pt, stats = serve(rate, n_req, seed)
        points.append(pt)
        if seed == 8 or reqs_sus is None:
            reqs_sus, pool_gib, shape_used = reqs, stats["pool_gib"], stats["tuning"]

    # FP32 served results against the FP64 form of the same request alone (one
    # request of each network size, its first check_trials trials)
    precision_check = None
    if precision == 1 and args.check_requests > 0:
        from unsupervised_growth_amd.engine import Engine
        tot = {"max_abs_dt_ms": 0.0, "unpaired_spikes": 0, "paired_within_one_step": 0, "spikes": 0,
               "requests": 0, "trials_per_request": 0}
        picked, seen = [], set()
        for r in reqs_sus:
            if r.net.n_ra not in seen and len(picked) < args.check_requests:
                seen.add(r.net.n_ra)
                picked.append(r)
        ntr = min(args.check_trials, args.trials)
        for r in picked:
            with Engine(precision=0, trial_ms=500.0, ra_model=args.ra_model, i_model=args.i_model,
                        seed=r.key) as e:
                e.upload(r.net)
                for k in range(ntr):
                    ref = e.run_trial(**r.trial_kw(k))
                    c = compare_results(ref, r.results[k], window=1.0)
                    tot["max_abs_dt_ms"] = max(tot["max_abs_dt_ms"], max(v["max_abs_dt_ms"] for v in c.values()))
                    tot["unpaired_spikes"] += sum(v["unpaired"] for v in c.values())
                    tot["paired_within_one_step"] += sum(v["paired_within_one_step"] for v in c.values())
                    tot["spikes"] += sum(v["spikes_ref"] for v in c.values())
            tot["requests"] += 1
        tot["trials_per_request"] = ntr
        tot["network_sizes"] = [r.net.n_ra for r in picked]
        precision_check = dict(reference="FP64 form of the same request run alone (bit-exact with the oracle)",
                               dt_ms=0.02, **tot)
    main_pt = points[-1]
    line = {
        "metric": "served chain-test requests: p50/p99 latency and neuron-updates/s",
        "value": main_pt["neuron_updates_per_s"], "unit": "neuron-updates/s", "n_gpus": 1,
        "steps": main_pt["requests"], "warmup": 0, "ms_per_step": 1000.0 * main_pt["wall_s"] / main_pt["requests"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == 0 else "f32",
        "data": "synthetic grown chain networks, Poisson arrivals (seeded)",
        "config": {"workload": "chain-test requests x %d trials of 500 ms, Poisson arrivals, N_RA 1000/4000/20000 "
                               "(25/50/25 %%), N_I = 0.275 N_RA, continuous batching on a %.1f GiB pool (%s)"
                               % (args.trials, pool_gib, "hvc_submit / hvc_poll" if args.serve == "async"
                                  else "rounds driven by the caller"),
                   "config": "c5", "network_build_s": round(t_build, 1)},
        "latency_s": main_pt["latency_s"],
        "requests_per_s": main_pt["requests_per_s"],
        "load_points": points,
        "tuning": shape_used,
    }
    if precision_check is not None:
        line["precision_check"] = precision_check
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build measures the single-GPU hot path "
              "(multi-GPU sharding is a design note, not an implemented path)", file=sys.stderr)
        sys.exit(2)
    if args.config == "c4":
        bench_batch(args)
        return
    if args.config == "c5":
        bench_serve(args)
        return
    if args.config == "c2":
        bench_growth_c2(args)
        return
    from unsupervised_growth_amd import configs, opcount
    from unsupervised_growth_amd.engine import Engine
    from unsupervised_growth_amd.spikes import compare_results

    t_build = time.perf_counter()
    net, kw = configs.CONFIGS[args.config]()
    t_build = time.perf_counter() - t_build
    if args.trial_ms:
        kw["trial_ms"] = args.trial_ms
    precision = 0 if args.precision == "fp64" else 1
    eng = Engine(ra_model=args.ra_model, i_model=args.i_model, precision=precision, **kw)
    eng.upload(net)
    tuning = apply_shape(eng, args)
    if args.copy_probe_mib > 0:
        # an elementwise kernel of known traffic (the engine's exp over N bytes of
        # doubles: reads and writes N bytes, N far above the L2 size): its byte
        # counters scale the kernels'
        import numpy as np
        eng.eval_math(0, np.zeros(args.copy_probe_mib << 17))
    steps_per_trial = int(round(kw["trial_ms"] / 0.02)) - 1

    def trial_kw(tr):
        return dict(mode=0, training=True, trial_index=tr, first_trial=(tr == 0), mean_injection_time=100.0)

    trial = 0
    for _ in range(args.warmup):
        eng.run_trial(**trial_kw(trial))
        trial += 1
    eng.set_timing(True)
    first_timed = trial
    wall = 0.0
    spikes = 0
    results, states = [], []
    for _ in range(args.steps):
        # the synapse state each trial starts from (outside the timed region), so
        # that the FP64 form can repeat the trial from the same state
        if precision == 1:
            states.append(eng.download_synapses())
        t0 = time.perf_counter()
        r = eng.run_trial(**trial_kw(trial))
        wall += time.perf_counter() - t0
        spikes += sum(map(len, r.soma)) + sum(map(len, r.interneuron))
        results.append(r)
        trial += 1
    ms_per_step = 1000.0 * wall / args.steps
    value = (net.n_ra + net.n_i) * steps_per_trial * args.steps / wall

    # roofline of the dominant kernel (HIP events on the engine stream)
    windows = 2000 if kw["trial_ms"] == 500.0 else int(round(kw["trial_ms"] / 0.25))
    kern = {}
    for name in ("integrate", "deliver", "sync", "decay"):
        avg, n = eng.kernel_timing(name)
        kern[name] = {"avg_ms": avg, "samples": n}
    eng.close()
    # integrate and deliver run once per window; sync (the last window's) and decay once per trial
    per_trial = {k: v["avg_ms"] * (1 if k in ("decay", "sync") else windows) for k, v in kern.items()}
    dom = max(per_trial, key=per_trial.get)
    ra_flops = opcount.flops_per_step("ra", args.ra_model, precision)
    i_flops = opcount.flops_per_step("i", args.i_model, precision)
    steps_per_window = steps_per_trial / windows
    flops_launch = (ra_flops * net.n_ra + i_flops * net.n_i) * steps_per_window
    avg_s = kern["integrate"]["avg_ms"] / 1000.0
    peak = FP64_PEAK_TFLOPS if precision == 0 else FP32_PEAK_TFLOPS
    achieved = flops_launch / avg_s / 1e12 if avg_s > 0 else 0.0
    traffic = TRAFFIC_PROFILED.get((args.config, precision, tuning["i_lanes"]))
    roofline = {"bound": "valu", "kernel": "integrate", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic, "unpacked_fp32": unpacked(achieved, precision),
                "traffic_note": None if traffic is None else "HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE "
                                "(scaled by a kernel of known traffic), profiles/integrate_c3_hbm_traffic.txt; "
                                "algorithmic state traffic ~7.5 MB per launch",
                "note": "vector-ALU bound (no dense contraction, no MFMA); peak = MI355X %s vector rate; achieved = "
                        "algorithmic flops per launch (opcount: %d per RA step, %d per I step, one exp = %d) / "
                        "HIP-event launch duration" % ("FP64" if precision == 0 else "FP32", ra_flops, i_flops,
                                                       opcount.FLOPS_PER_EXP[precision]),
                "flops_per_launch": flops_launch, "avg_launch_ms": kern["integrate"]["avg_ms"],
                "dominant": dom, "kernels_ms_per_trial": per_trial}
    deliver = deliver_bytes(net, results, windows, kern["deliver"]["avg_ms"], (args.config, tuning["deliver_lists"]))

    # FP32 form: each timed trial again in the FP64 form (bit-exact with the
    # oracle) from the same synapse state; per-trial spike-time error
    precision_check = None
    fp64_form = None
    if precision == 1:
        per = []
        wall64 = 0.0
        with Engine(ra_model=args.ra_model, i_model=args.i_model, precision=0, **kw) as e64:
            e64.upload(net)
            tuning64 = apply_shape(e64, args)
            for k, syn in enumerate(states):
                e64.upload_synapses(syn)
                t0 = time.perf_counter()
                r64 = e64.run_trial(**trial_kw(first_timed + k))
                wall64 += time.perf_counter() - t0
                c = compare_results(r64, results[k], window=1.0)
                per.append({"trial": first_timed + k,
                            "max_abs_dt_ms": max(v["max_abs_dt_ms"] for v in c.values()),
                            "unpaired_spikes": sum(v["unpaired"] for v in c.values()),
                            "paired_within_one_step": sum(v["paired_within_one_step"] for v in c.values()),
                            "spikes": sum(v["spikes_ref"] for v in c.values())})
        precision_check = {"reference": "FP64 form (bit-exact with the oracle), each trial from the same synapse state",
                           "dt_ms": 0.02, "max_abs_dt_ms": max(p["max_abs_dt_ms"] for p in per),
                           "unpaired_spikes": sum(p["unpaired_spikes"] for p in per), "per_trial": per}
        fp64_form = {"value": (net.n_ra + net.n_i) * steps_per_trial * len(states) / wall64,
                     "ms_per_step": 1000.0 * wall64 / len(states),
                     "tuning": tuning64,
                     "note": "the same trials in the plain FP64 GPU form (bit-exact with the oracle)"}
    growth_loop = bench_driver(net, kw, args, precision, tuning) if (args.config == "c3" and
                                                                      args.driver_trials > 0) else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(net, kw, args)
    line = {
        "metric": "RK4 neuron-updates/s at 20k RA + 5.5k I; p50/p99 served trial latency",
        "value": value, "unit": "neuron-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == 0 else "f32",
        "data": "synthetic network (distance rules, seeded), counter-based noise streams",
        "config": {"workload": "growth trial, %d HVC-RA + %d HVC-I, %d RA->RA candidates, %g ms trials, "
                               "RA %s + I %s, STDP on, %s state" % (
                                   net.n_ra, net.n_i, net.n_candidates, kw["trial_ms"],
                                   "RK4-3/8" if args.ra_model else "Debraband",
                                   "DP8" if args.i_model == 0 else "RK4", "FP64" if precision == 0 else "FP32"),
                   "config": args.config, "network_build_s": round(t_build, 1),
                   "spikes_per_trial": spikes / args.steps, "tuning": tuning},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line["deliver"] = deliver
    if precision_check is not None:
        line["precision_check"] = precision_check
        line["fp64_form"] = fp64_form
    if args.dropin_trials > 0:
        line["dropin"] = bench_dropin(net, kw, args, tuning)
    if growth_loop is not None:
        line["growth_loop"] = growth_loop
    print(json.dumps(line), flush=True)

if __name__ == "__main__":
    main()
