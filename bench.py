"""Benchmark: particle-steps/s of the SPH step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload served_stream]

Default workload: the served scene stream -- seeded requests (50% fluid,
30% corotational, 20% SVK blocks, 1k-262k particles log-uniform, K = 200
steps each) run through the device batch (continuous batching, hipGraph
step replay, HBM-budget admission) and, for comparison, one scene at a
time through the single-scene path.  `value` is the batch's throughput with
every request queued at the start (saturated); the same line reports the
single-scene path on a prefix of the stream, the batch on that prefix, and
p50/p99 request latency with Poisson arrivals at --rate per second.
Scenes, results and precision settings are identical between the two paths
(bitwise, as the GPU tests check), so the comparison is at fixed error.

Single-scene workloads (--workload dam_break_4m etc.): a step is one full
simulation step (neighbour grid, density, forces, integrator) over the whole
scene with its state resident in HBM.

Each line also carries the roofline of the dominant kernel (HIP-event timing
over a profiled pass) and a CPU baseline (the FP32 oracle on the host cores,
bounded sample).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "opencl-sph_amd"))

METRIC = "particle-steps/s batched vs plain at fixed FP64-oracle error; p50/p99 latency"
WORKLOADS = {"dam_break_16k": 0, "cantilever_16k": 1, "svk_cube_262k": 2, "dam_break_4m": 3, "served_stream": 4}
SUMS = {"plain": 0, "kahan-density": 1, "kahan-all": 3}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def grid_pair_counts(system, scene):
    """Exact candidate and in-range pair counts of the current grid."""
    import numpy as np
    system.sync_fields_device_to_host("gridcount", "num_neighbours")
    count, nn = system.get_fields("gridcount", "num_neighbours")
    c = np.asarray(count, np.int64)               # (rz, ry, rx)
    pad = np.pad(c, 1)
    box = np.zeros_like(c)
    rz, ry, rx = c.shape
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                box += pad[dz:dz + rz, dy:dy + ry, dx:dx + rx]
    candidates = int((c * box).sum())
    in_range = int(np.asarray(nn[:scene.n], np.int64).sum())
    return candidates, in_range


# Timed pass name -> device symbol prefix in the rocprofv3 counter summaries
# (a pass may launch several kernels with that prefix: their traffic adds up).
KERNEL_SYMBOLS = {"compute_density": "k_density_listed<sphk::SingleScene",
                  "compute_forces_fluids": "k_forces_fluids_listed<sphk::SingleScene",
                  "bin_rank_place": "k_rank_place", "step_forward": "k_step_forward<sphk::SingleScene",
                  "bin_scatter": "k_scatter_unsorted", "bin_key_count": "k_key_count"}
BATCH_KERNEL_SYMBOLS = {
    "compute_density_fluid": "k_density_listed<sphk::SceneTable",
    "compute_forces_fluids": "k_forces_fluids_listed<sphk::SceneTable",
    "compute_density_solid": "k_density_solid_listed<sphk::SceneTable",
    "compute_rotations_and_strains": "k_rotations_strains<sphk::SceneTable",
    "compute_deformations": "k_deformations_svk<sphk::SceneTable",
    "compute_forces_solids_corotational": "k_forces_solids<sphk::SceneTable, true",
    "compute_forces_solids_svk": "k_forces_solids<sphk::SceneTable, false",
    "step_forward": "k_step_forward<sphk::SceneTable", "batch_rank_gather": "k_batch_rank_gather",
}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of pass `kernel` from the calibrated FETCH_SIZE /
    WRITE_SIZE summary committed under profiles/, or None."""
    path = os.path.join(REPO, "profiles", f"{workload}_hbm_traffic.json")
    table = BATCH_KERNEL_SYMBOLS if workload == "served_stream" else KERNEL_SYMBOLS
    prefix = table.get(kernel)
    if prefix is None or not os.path.exists(path):
        return None
    with open(path) as f:
        summary = json.load(f)
    rows = [row for name, row in summary.get("kernels", {}).items() if name.startswith(prefix)]
    return int(sum(r["read_bytes"] + r["write_bytes"] for r in rows)) if rows else None


def measured_issue(workload, kernel):
    """Share of CU cycles in which pass `kernel` issued a wave64 VALU
    instruction (rocprofv3 SQ counters, summary committed under profiles/),
    or None.  The neighbour passes are bound by that issue rate, not by HBM."""
    path = os.path.join(REPO, "profiles", "neighbour_passes_instruction_mix.json")
    table = BATCH_KERNEL_SYMBOLS if workload == "served_stream" else KERNEL_SYMBOLS
    prefix = table.get(kernel)
    if prefix is None or not os.path.exists(path):
        return None
    with open(path) as f:
        rows = [r for name, r in json.load(f).get(workload, {}).items() if name.startswith(prefix)]
    # kernels of one pass (list and overflow launches): issued VALU over cycles
    valu = sum(r["valu_per_wave"] * r["waves"] for r in rows)
    cycles = sum(r["valu_per_wave"] * r["waves"] / r["valu_issue"] for r in rows if r["valu_issue"] > 0)
    return round(valu / cycles, 3) if cycles else None


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB (a copy kernel of known size):
    profiled beside the workload, its FETCH_SIZE / WRITE_SIZE readings give
    the scale that converts those counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed in the copy probe")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: 3 x {mib} MiB device-to-device")


def cpu_baseline(scenes_, budget_s):
    """The FP32 oracle on the host cores over the given scenes (reference
    configuration set up for solids, one warm-up step each), as many whole
    steps as fit an equal share of the time budget."""
    import ctypes

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    lib, _, _ = oracle.load("f32")
    lib.orc_threads.restype = ctypes.c_uint
    threads = int(lib.orc_threads())
    work = elapsed = 0.0
    parts = []
    for scene in scenes_:
        fields = {k: v for k, v in scene.fields.items() if k in oracle.OracleState.PER_PARTICLE}
        for k in ("body_number", "body_dynamic"):
            if k in scene.fields:
                fields[k] = scene.fields[k]
        params = dict(scene.params)
        params["polar_fp64"] = 1
        st = oracle.OracleState("f32", params, **fields)
        if scene.system in ("corotational", "svk"):
            st.setup_solid()
        st.step(scene.system, 1)                    # warm-up
        steps, t0 = 0, time.perf_counter()
        while True:
            st.step(scene.system, 1)
            steps += 1
            el = time.perf_counter() - t0
            if el >= budget_s / len(scenes_) or steps >= scene.steps:
                break
        work += scene.n * steps
        elapsed += el
        parts.append(f"{steps} step(s) of {scene.name} ({scene.system}, {scene.n} particles)")
    return {"value": work / elapsed, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "; ".join(parts) + f"; FP32 oracle (C restatement of the reference kernels, OpenMP), "
                      f"after 1 warm-up step each, {elapsed:.1f} s wall"}


def run_single(args):
    import numpy as np

    import sph
    from sph import autotune, roofline, scenes

    scene = scenes.baseline_config(WORKLOADS[args.workload])
    if args.autotune == "run":
        tuning, _ = autotune.tune_single(scene, log=log)
    elif args.autotune == "load":
        tuning = autotune.lookup(autotune.single_signature(scene))
    else:
        tuning = dict(autotune.DEFAULTS)
    autotune.apply(tuning)
    log(f"workload {scene.name}: {scene.n} particles, system {scene.system}")
    system = scenes.build(scene)
    handle = system.psdata_opencl
    handle.set_precision(compensated_sums=SUMS[args.sums], polar_fp64=True)

    system.run(args.warmup)
    handle.synchronize()
    log(f"warm-up {args.warmup} steps done")

    t0 = time.perf_counter()
    system.run(args.steps)
    handle.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = scene.n * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.4g} particle-steps/s")

    handle.reset_kernel_stats()
    handle.set_profiling(True)
    system.run(args.steps)
    handle.synchronize()
    stats = handle.kernel_stats()
    handle.set_profiling(False)
    per_step = {k: ms / args.steps for k, (launches, ms) in stats.items()}
    log("per-kernel ms/step: " + ", ".join(f"{k}={v:.4f}" for k, v in sorted(per_step.items(), key=lambda x: -x[1])))

    cells = int(np.prod(system.get_fields("gridcount")[0].shape))
    dominant = max(stats, key=lambda k: stats[k][1])
    launches, total_ms = stats[dominant]
    avg_s = total_ms / launches / 1e3
    bytes_launch = roofline.bytes_per_launch(dominant, scene.n, cells)
    roof = None
    if bytes_launch is not None:
        achieved = bytes_launch / avg_s / 1e9
        roof = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": roofline.MI355X_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / roofline.MI355X_HBM_GBS, 4),
                "traffic": measured_traffic(scene.name, dominant),
                "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg_s * 1e3, 5),
                "valu_issue": measured_issue(scene.name, dominant)}
    extra = {}
    if scene.system == "fluid":
        candidates, in_range = grid_pair_counts(system, scene)
        flops = roofline.flops_per_launch(dominant, candidates, in_range)
        if roof is not None and flops:
            roof["valu_tflops"] = round(flops / avg_s / 1e12, 3)
            roof["valu_frac"] = round(flops / avg_s / 1e12 / roofline.MI355X_FP32_TFLOPS, 4)
        t_roof = roofline.t_roof_seconds(scene.n, cells, candidates, in_range)
        # the streaming (HBM-bound) kernels against the HBM peak, each on its
        # own algorithmic bytes
        streaming = {}
        for k in ("bin_key_count", "bin_scan", "bin_scatter", "bin_rank_place", "step_forward"):
            if k not in stats:
                continue
            n_launch, ms = stats[k]
            b = roofline.bytes_per_launch(k, scene.n, cells)
            gbs = b / (ms / n_launch / 1e3) / 1e9
            streaming[k] = {"bytes_per_launch": b, "avg_launch_ms": round(ms / n_launch, 5),
                            "achieved_gbs": round(gbs, 1), "frac": round(gbs / roofline.MI355X_HBM_GBS, 4)}
        extra = {"step_roofline": {"t_roof_ms": round(t_roof * 1e3, 5), "frac": round(t_roof * 1e3 / ms_per_step, 4),
                                   "algorithmic_bytes_per_step": roofline.step_bytes(scene.n, cells),
                                   "candidate_pairs": candidates, "in_range_pairs": in_range},
                 "hbm_kernels": streaming,
                 "kernel_ms_per_step": {k: round(v, 5) for k, v in per_step.items()}}
    system.free()

    cpu = None
    if args.cpu_budget > 0:
        log("CPU baseline (FP32 oracle) ...")
        cpu = cpu_baseline([scene], args.cpu_budget)
        log(f"CPU baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded lattice with uniform jitter)",
        "config": {"workload": scene.name, "particles": scene.n, "cells": cells, "system": scene.system,
                   "path": "plain", "sums": args.sums, "tuning": tuning,
                   "fluid_model": "tait eos, mass-consistent force" if scene.system == "fluid" else None},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(extra)
    return line


def _brief(rep):
    keep = ("requests", "particle_steps", "wall_s", "particle_steps_per_s", "latency_p50_s", "latency_p99_s",
            "latency_mean_s", "latency_max_s", "peak_resident_slots", "peak_resident_scenes", "batch_steps",
            "graph_captures", "graph_launches", "pool_bytes", "host_phases_s")
    return {k: (round(v, 5) if isinstance(v, float) else v) for k, v in rep.items() if k in keep}


def run_served(args):
    from sph import autotune, roofline, serve

    steps = args.steps if args.steps_set else 200
    if args.autotune == "run":
        # tuned on requests of another seed, not on the measured stream
        sample = serve.make_requests(count=64, rate=args.rate, seed=args.seed + 2, steps=60)
        tuning, _ = autotune.tune_batch(sample, log=log)
    elif args.autotune == "load":
        tuning = autotune.lookup(autotune.batch_signature())
    else:
        tuning = dict(autotune.DEFAULTS)
    autotune.apply(tuning)
    if args.graph_steps is None:
        args.graph_steps = tuning["graph_steps"]
    sums = SUMS[args.sums]
    reqs = serve.make_requests(count=args.requests, rate=args.rate, seed=args.seed, steps=steps)
    prefix = reqs[:args.plain_requests]
    budget = args.hbm_budget_gb * 1e9
    log(f"served stream: {len(reqs)} requests, {steps} steps each, "
        f"{sum(r.n * r.steps for r in reqs):.4g} particle-steps, last arrival {reqs[-1].arrival:.1f} s")

    if args.counters_only:
        # Only the profiled pass, for hardware-counter runs under rocprofv3,
        # after a batched warm-up like the one of the full run, so that the
        # traced pass meets the device in the same state as the bench's own.
        warm = serve.make_requests(count=max(args.warmup, 1), rate=args.rate, seed=args.seed + 1, steps=4)
        serve.serve_batched(warm, hbm_budget=budget, realtime=False, compensated_sums=sums,
                            graph_steps=args.graph_steps)
        prof = serve.serve_batched(reqs[:args.profile_requests], hbm_budget=budget, realtime=False,
                                   compensated_sums=sums, profile=True)
        return {"counters_only": True, "requests": args.profile_requests,
                "particle_steps_by_system": prof["particle_steps_by_system"],
                "kernel_stats": prof["kernel_stats"]}

    # warm-up: both paths on a few requests of another seed
    warm = serve.make_requests(count=max(args.warmup, 1), rate=args.rate, seed=args.seed + 1, steps=4)
    serve.serve_batched(warm, hbm_budget=budget, realtime=False, compensated_sums=sums, graph_steps=args.graph_steps)
    serve.serve_plain(warm[:2], realtime=False, compensated_sums=sums)
    log("warm-up done")

    sat = serve.serve_batched(reqs, hbm_budget=budget, realtime=False, compensated_sums=sums,
                              graph_steps=args.graph_steps, log=log)
    log(f"batched, saturated: {sat['particle_steps_per_s']:.4g} particle-steps/s over {sat['wall_s']:.1f} s")
    sat_prefix = serve.serve_batched(prefix, hbm_budget=budget, realtime=False, compensated_sums=sums,
                                     graph_steps=args.graph_steps)
    plain_prefix = serve.serve_plain(prefix, realtime=False, compensated_sums=sums, log=log)
    log(f"first {len(prefix)} requests, saturated: batched {sat_prefix['particle_steps_per_s']:.4g}, "
        f"plain {plain_prefix['particle_steps_per_s']:.4g} particle-steps/s")
    live = serve.serve_batched(reqs, hbm_budget=budget, realtime=True, compensated_sums=sums,
                               graph_steps=args.graph_steps, log=log)
    log(f"batched, Poisson {args.rate}/s: p50 {live['latency_p50_s'] * 1e3:.1f} ms, "
        f"p99 {live['latency_p99_s'] * 1e3:.1f} ms")
    live_plain = None
    if args.plain_served:
        live_plain = serve.serve_plain(prefix, realtime=True, compensated_sums=sums, log=log)
        log(f"plain, Poisson {args.rate}/s, first {len(prefix)}: p50 {live_plain['latency_p50_s'] * 1e3:.1f} ms, "
            f"p99 {live_plain['latency_p99_s'] * 1e3:.1f} ms")

    # profiled pass: per-pass HIP-event times on a prefix, without graphs
    prof = serve.serve_batched(reqs[:args.profile_requests], hbm_budget=budget, realtime=False,
                               compensated_sums=sums, profile=True)
    stats = prof["kernel_stats"]
    dominant = max(stats, key=lambda k: stats[k][1])
    launches, total_ms = stats[dominant]
    avg_s = total_ms / launches / 1e3
    bytes_launch = roofline.batch_bytes_per_launch(dominant, prof["particle_steps_by_system"], launches)
    roof = None
    if bytes_launch:
        achieved = bytes_launch / avg_s / 1e9
        roof = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": roofline.MI355X_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / roofline.MI355X_HBM_GBS, 5),
                "traffic": measured_traffic("served_stream", dominant),
                "valu_issue": measured_issue("served_stream", dominant),
                "bytes_per_launch": round(bytes_launch), "avg_launch_ms": round(avg_s * 1e3, 5),
                "sample": f"first {len(reqs[:args.profile_requests])} requests, saturated, no graphs"}
    pass_ms = {k: round(ms / max(n, 1), 5) for k, (n, ms) in sorted(stats.items(), key=lambda x: -x[1][1])}
    log("per-pass ms/launch: " + ", ".join(f"{k}={v}" for k, v in pass_ms.items()))

    cpu = None
    if args.cpu_budget > 0:
        log("CPU baseline (FP32 oracle) ...")
        picks = []
        for kind in serve.KINDS:
            r = next(q for q in reqs if q.kind == kind)
            picks.append(r.scene())
        cpu = cpu_baseline(picks, args.cpu_budget)
        log(f"CPU baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} threads")

    return {
        "metric": METRIC, "value": round(sat["particle_steps_per_s"], 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(sat["wall_s"] / max(sat["batch_steps"], 1) * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (seeded request stream, seed {args.seed}: blocks with seeded jitter / smooth velocity)",
        "config": {"workload": "served_stream", "requests": len(reqs), "rate_per_s": args.rate,
                   "mix": "50% fluid / 30% corotational / 20% svk", "particles": "log-uniform 1024-262144",
                   "steps_per_request": steps, "hbm_budget_gb": args.hbm_budget_gb, "sums": args.sums,
                   "polar": "fp64", "graph_steps": args.graph_steps, "pools": sat.get("pools", 1), "path": "batched",
                   "tuning": tuning},
        "batched_saturated": _brief(sat),
        "prefix": {"requests": len(prefix), "batched": _brief(sat_prefix), "plain": _brief(plain_prefix),
                   "speedup": round(sat_prefix["particle_steps_per_s"] / plain_prefix["particle_steps_per_s"], 3)},
        "served": {"batched": _brief(live), "plain_prefix": _brief(live_plain) if live_plain else None},
        "error_vs_single_scene": "bitwise identical (tests/test_gpu_batch.py)",
        "roofline": roof, "pass_ms_per_launch": pass_ms, "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="steps per request (served stream, default 200) or timed steps (single scene, default 50)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="served_stream", choices=sorted(WORKLOADS))
    ap.add_argument("--sums", default="kahan-density", choices=sorted(SUMS),
                    help="neighbour-sum precision (default: Kahan-compensated density sums)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-baseline work (0: skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a device copy of this many MiB (counter calibration under rocprofv3)")
    ap.add_argument("--requests", type=int, default=2048, help="served stream: number of requests")
    ap.add_argument("--rate", type=float, default=40.0, help="served stream: Poisson arrival rate per second")
    ap.add_argument("--seed", type=int, default=0, help="served stream: seed")
    ap.add_argument("--plain-requests", type=int, default=192,
                    help="served stream: prefix run through the single-scene path for comparison")
    ap.add_argument("--plain-served", type=int, default=1, help="served stream: also serve the prefix one at a time")
    ap.add_argument("--profile-requests", type=int, default=64, help="served stream: prefix of the profiled pass")
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0, help="served stream: admission budget")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="served stream: batch steps per hipGraph (even; default: tuned)")
    ap.add_argument("--autotune", default="load", choices=["load", "run", "off"],
                    help="launch parameters: the tuned table (load), tune now (run), or defaults (off)")
    ap.add_argument("--counters-only", action="store_true",
                    help="served stream: run only the profiled pass (for rocprofv3 counter collection)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the engine runs one scene set per device and this build has no "
              f"multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    args.steps_set = args.steps is not None
    if args.steps is None:
        args.steps = 50

    import sph
    from sph import bindings

    bindings.set_log_level(3)          # library notes would share stdout with the JSON line
    sph.init_hip()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
    line = run_served(args) if args.workload == "served_stream" else run_single(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
