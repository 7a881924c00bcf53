"""Benchmark: particle-steps/s of the SPH step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload served_stream]

Default workload: the served scene stream -- 2048 seeded requests (50%
fluid, 30% corotational, 20% SVK blocks, 1k-262k particles log-uniform, 200
steps each, --request-steps) run through the device batch (continuous
batching, hipGraph step replay, HBM-budget admission) and, for comparison,
one scene at a time through the single-scene path.  `value` is the batch's
throughput with every request queued at the start (saturated); a "step" of
the line is one batch step (every resident scene advances one step), so
`steps` x `ms_per_step` is the saturated run's wall time.  --steps and
--warmup do not change the stream: --warmup is the number of warm-up
requests (of another seed).  The same line reports the single-scene path
on a prefix of the stream, the batch on that prefix, p50/p99 request
latency with Poisson arrivals at --rate per second, and the 4.19M-particle
dam break (50 steps, plain single-scene path) with its step roofline.
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
                  "compute_forces_fluids+step_forward": "k_forces_fluids_listed<sphk::SingleScene",
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


def host_description():
    """nproc, the CPU set this process may use, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def cpu_baseline(scenes_, budget_s, compensated=1, repeats=5):
    """The FP32 oracle on the host cores, per system: after one warm-up step,
    `repeats` timed repeats of k steps each (k sized so that all repeats of
    all systems fit `budget_s`); the median repeat gives the system's rate.
    The overall value is total particle-steps over the summed medians."""
    import ctypes
    import statistics

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    lib, _, _ = oracle.load("f32")
    lib.orc_threads.restype = ctypes.c_uint
    threads = int(lib.orc_threads())
    per_system = {}
    work = med_total = 0.0
    share = budget_s / max(len(scenes_), 1) / repeats
    for scene in scenes_:
        fields = {k: v for k, v in scene.fields.items() if k in oracle.OracleState.PER_PARTICLE}
        for k in ("body_number", "body_dynamic"):
            if k in scene.fields:
                fields[k] = scene.fields[k]
        params = dict(scene.params)
        params["polar_fp64"] = 1
        params["compensated"] = compensated
        st = oracle.OracleState("f32", params, **fields)
        if scene.system in ("corotational", "svk"):
            st.setup_solid()
        t0 = time.perf_counter()
        st.step(scene.system, 1)                    # warm-up, also sizes a repeat
        one = time.perf_counter() - t0
        k = max(1, min(int(share / max(one, 1e-6)), scene.steps))
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            st.step(scene.system, k)
            times.append(time.perf_counter() - t0)
        med = statistics.median(times)
        per_system[scene.system] = {"value": scene.n * k / med, "scene": scene.name, "particles": scene.n,
                                    "steps_per_repeat": k, "median_s": round(med, 4),
                                    "repeats_s": [round(t, 4) for t in times]}
        work += scene.n * k
        med_total += med
    host = host_description()
    return {"value": work / med_total, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "; ".join(f"{v['scene']} ({k}, {v['particles']} particles): {v['steps_per_repeat']} steps x "
                                f"{repeats} repeats, median {v['median_s']} s" for k, v in per_system.items())
                      + f"; FP32 oracle (C restatement of the reference kernels, OpenMP, compensated sums "
                        f"{compensated}, FP64 polar), after 1 warm-up step",
            "per_system": per_system, "repeats": repeats, "threads": threads, **host,
            "reference_cpu_opencl": "unavailable: the reference creates a GPU-only OpenCL context "
                                    "(src/opencl/particle_system_host.c:65-88) and its solid programs do not "
                                    "compile under clang-based OpenCL; no CPU OpenCL device is installed"}


def pass_roofline(name, n_launch, ms, nbytes, flops, bound):
    """Roofline entry of one pass from its total time over n_launch launches
    and its algorithmic bytes / flops per launch.  `bound` names the peak the
    fraction is taken against ("hbm" or "valu"); both rates are reported."""
    from sph import roofline
    avg_s = ms / n_launch / 1e3
    out = {"avg_launch_ms": round(avg_s * 1e3, 5), "bound": bound}
    if nbytes is not None:
        out["bytes_per_launch"] = int(nbytes)
        out["achieved_gbs"] = round(nbytes / avg_s / 1e9, 2)
        out["hbm_frac"] = round(nbytes / avg_s / 1e9 / roofline.MI355X_HBM_GBS, 5)
    if flops:
        out["flops_per_launch"] = int(flops)
        out["achieved_tflops"] = round(flops / avg_s / 1e12, 3)
        out["valu_frac"] = round(flops / avg_s / 1e12 / roofline.MI355X_FP32_TFLOPS, 5)
    out["frac"] = out.get("valu_frac" if bound == "valu" else "hbm_frac")
    return out


def roofline_block(kernel, entry, traffic, issue, sample=None):
    """The line's `roofline` object for the dominant kernel."""
    from sph import roofline
    valu = entry["bound"] == "valu"
    block = {"kernel": kernel, "bound": entry["bound"],
             "achieved": entry["achieved_tflops"] if valu else entry["achieved_gbs"],
             "peak": roofline.MI355X_FP32_TFLOPS if valu else roofline.MI355X_HBM_GBS,
             "unit": "TFLOP/s" if valu else "GB/s", "frac": entry["frac"], "traffic": traffic,
             "avg_launch_ms": entry["avg_launch_ms"], "bytes_per_launch": entry.get("bytes_per_launch"),
             "flops_per_launch": entry.get("flops_per_launch"), "hbm_frac": entry.get("hbm_frac"),
             "valu_frac": entry.get("valu_frac"), "valu_issue": issue}
    if sample:
        block["sample"] = sample
    return block


def measure_single(scene, steps, warmup, sums, tuning):
    """K timed steps of one scene on the single-scene path (state resident in
    HBM), then K profiled steps with per-pass HIP-event timing; returns the
    measurements and, for fluids, the step roofline with every pass's
    fraction of the bound that applies to it."""
    import numpy as np

    from sph import autotune, roofline, scenes

    autotune.apply(tuning)
    log(f"workload {scene.name}: {scene.n} particles, system {scene.system}")
    system = scenes.build(scene)
    handle = system.psdata_opencl
    handle.set_precision(compensated_sums=SUMS[sums], polar_fp64=True)

    system.run(warmup)
    handle.synchronize()
    log(f"warm-up {warmup} steps done")
    t0 = time.perf_counter()
    system.run(steps)
    handle.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / steps * 1e3
    value = scene.n * steps / elapsed
    log(f"timed {steps} steps: {ms_per_step:.3f} ms/step, {value:.4g} particle-steps/s")

    handle.reset_kernel_stats()
    handle.set_profiling(True)
    system.run(steps)
    handle.synchronize()
    stats = handle.kernel_stats()
    handle.set_profiling(False)
    per_step = {k: ms / steps for k, (launches, ms) in stats.items()}
    log("per-kernel ms/step: " + ", ".join(f"{k}={v:.4f}" for k, v in sorted(per_step.items(), key=lambda x: -x[1])))
    cells = int(np.prod(system.get_fields("gridcount")[0].shape))
    out = {"scene": scene, "value": value, "ms_per_step": ms_per_step, "steps": steps, "cells": cells,
           "stats": stats, "per_step": per_step}
    if scene.system == "fluid":
        candidates, in_range = grid_pair_counts(system, scene)
        passes = {}
        for k, (n_launch, ms) in stats.items():
            nbytes = roofline.bytes_per_launch(k, scene.n, cells)
            flops = roofline.flops_per_launch(k, candidates, in_range)
            if nbytes is None:
                continue
            passes[k] = pass_roofline(k, n_launch, ms, nbytes, flops, "valu" if flops else "hbm")
        t_roof = roofline.t_roof_seconds(scene.n, cells, candidates, in_range)
        out["passes"] = passes
        out["step_roofline"] = {"t_roof_ms": round(t_roof * 1e3, 5), "frac": round(t_roof * 1e3 / ms_per_step, 4),
                                "algorithmic_bytes_per_step": roofline.step_bytes(scene.n, cells),
                                "candidate_pairs": candidates, "in_range_pairs": in_range,
                                "rule": "sum over passes of max(bytes / 8 TB/s, flops / 157.3 TFLOP/s)"}
    system.free()
    return out


def run_single(args):
    from sph import autotune, scenes

    scene = scenes.baseline_config(WORKLOADS[args.workload])
    if args.autotune == "run":
        tuning, _ = autotune.tune_single(scene, log=log)
    elif args.autotune == "load":
        tuning = autotune.lookup(autotune.single_signature(scene))
    else:
        tuning = dict(autotune.DEFAULTS)
    m = measure_single(scene, args.steps, args.warmup, args.sums, tuning)
    stats = m["stats"]
    dominant = max(stats, key=lambda k: stats[k][1])
    roof = None
    if dominant in m.get("passes", {}):
        roof = roofline_block(dominant, m["passes"][dominant], measured_traffic(scene.name, dominant),
                              measured_issue(scene.name, dominant))
    cpu = None
    if args.cpu_budget > 0:
        log("CPU baseline (FP32 oracle) ...")
        cpu = cpu_baseline([scene], args.cpu_budget, SUMS[args.sums])
        log(f"CPU baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} threads")
    line = {
        "metric": METRIC, "value": round(m["value"], 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m["ms_per_step"], 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded lattice with uniform jitter)",
        "config": {"workload": scene.name, "particles": scene.n, "cells": m["cells"], "system": scene.system,
                   "path": "plain", "sums": args.sums, "tuning": tuning,
                   "fluid_model": "tait eos, mass-consistent force" if scene.system == "fluid" else None},
        "roofline": roof, "cpu_baseline": cpu,
        "kernel_ms_per_step": {k: round(v, 5) for k, v in m["per_step"].items()},
    }
    if "step_roofline" in m:
        line["step_roofline"] = m["step_roofline"]
        line["passes"] = m["passes"]
    return line


def dam_break_4m_block(args):
    """The 4.19M-particle dam break on the plain single-scene path, 50 timed
    steps (its own step count), with the step roofline and every pass's
    fraction of its bound."""
    from sph import autotune, scenes
    scene = scenes.baseline_config(3)
    tuning = autotune.lookup(autotune.single_signature(scene)) if args.autotune != "off" else dict(autotune.DEFAULTS)
    m = measure_single(scene, scene.steps, 5, args.sums, tuning)
    return {"workload": scene.name, "particles": scene.n, "cells": m["cells"], "steps": m["steps"],
            "ms_per_step": round(m["ms_per_step"], 5), "particle_steps_per_s": round(m["value"], 1),
            "path": "plain (single scene)", "sums": args.sums, "tuning": tuning,
            "step_roofline": m["step_roofline"], "passes": m["passes"],
            "traffic_per_launch": {k: measured_traffic(scene.name, k) for k in m["passes"]}}


def _brief(rep):
    keep = ("requests", "particle_steps", "wall_s", "particle_steps_per_s", "latency_p50_s", "latency_p99_s",
            "latency_mean_s", "latency_max_s", "peak_resident_slots", "peak_resident_scenes", "batch_steps",
            "graph_captures", "graph_launches", "pool_bytes", "host_phases_s", "gpu_busy_lower_bound",
            "prepare_wait_share")
    return {k: (round(v, 5) if isinstance(v, float) else v) for k, v in rep.items() if k in keep}


def scene_pair_counts(scene):
    """(candidate pairs, in-range pairs with length <= h, self included) of
    a scene's initial state: candidates from the 27 cells around each
    particle's cell of the reference's grid (cell edge |hi - lo| / res >= h),
    in-range pairs from a k-d tree.  Solids move by a small fraction of h
    over a request, so these are also the pairs of each of its steps."""
    import numpy as np
    from scipy.spatial import cKDTree

    pos = np.asarray(scene.fields["position"], np.float32)[:scene.n]
    gb = np.asarray(scene.params["gridbounds"], np.float32).reshape(3, 2)
    h = np.float32(scene.params["h"])
    res = np.array([int(abs(float(gb[a, 1] - gb[a, 0])) / float(h)) for a in range(3)])
    v = (pos - gb[:, 0]) / (gb[:, 1] - gb[:, 0]) * res.astype(np.float32)
    cell = np.clip(np.where(v > 0, v, 0).astype(np.int64), 0, res - 1)
    key = (cell[:, 2] * res[1] + cell[:, 1]) * res[0] + cell[:, 0]
    count = np.bincount(key, minlength=int(np.prod(res))).reshape(res[2], res[1], res[0])
    pad = np.pad(count, 1)
    box = np.zeros_like(count)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                box += pad[dz:dz + res[2], dy:dy + res[1], dx:dx + res[0]]
    candidates = int((count * box).sum())
    tree = cKDTree(pos.astype(np.float64))
    in_range = int(tree.count_neighbors(tree, float(h)))
    return candidates, in_range


def run_served(args):
    from sph import autotune, roofline, serve

    steps = args.request_steps
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
    warm = serve.make_requests(count=max(args.warmup, 1), rate=args.rate, seed=args.seed + 1, steps=4)

    if args.counters_only:
        # Only the profiled pass, for hardware-counter runs under rocprofv3,
        # after a batched warm-up like the one of the full run, so that the
        # traced pass meets the device in the same state as the bench's own.
        serve.serve_batched(warm, hbm_budget=budget, realtime=False, compensated_sums=sums,
                            graph_steps=args.graph_steps)
        prof = serve.serve_batched(reqs[:args.profile_requests], hbm_budget=budget, realtime=False,
                                   compensated_sums=sums, profile=True)
        return {"counters_only": True, "requests": args.profile_requests,
                "particle_steps_by_system": prof["particle_steps_by_system"],
                "kernel_stats": prof["kernel_stats"]}

    # warm-up: both paths on a few requests of another seed
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
    sample = reqs[:args.profile_requests]
    prof = serve.serve_batched(sample, hbm_budget=budget, realtime=False, compensated_sums=sums, profile=True)
    stats = prof["kernel_stats"]
    pair_steps = {}
    for r in sample:
        c, p = scene_pair_counts(r.scene())
        a, b = pair_steps.get(r.kind, (0, 0))
        pair_steps[r.kind] = (a + c * r.steps, b + p * r.steps)
    passes = {}
    for k, (n_launch, ms) in stats.items():
        if not n_launch:
            continue
        nbytes = roofline.batch_bytes_per_launch(k, prof["particle_steps_by_system"], n_launch)
        flops = roofline.batch_flops(k, pair_steps, bool(sums & 1))
        passes[k] = pass_roofline(k, n_launch, ms, nbytes, flops / n_launch if flops else None,
                                  "valu" if flops else "hbm")
    dominant = max(stats, key=lambda k: stats[k][1])
    roof = roofline_block(dominant, passes[dominant], measured_traffic("served_stream", dominant),
                          measured_issue("served_stream", dominant),
                          sample=f"first {len(sample)} requests ({steps} steps each), saturated, no graphs; "
                                 f"pair counts of the requests' initial states")
    log("per-pass ms/launch: " + ", ".join(f"{k}={v['avg_launch_ms']}" for k, v in passes.items()))

    big = None
    if args.dam_break_4m:
        log("4.19M-particle dam break, plain path ...")
        big = dam_break_4m_block(args)
        log(f"4.19M dam break: {big['ms_per_step']:.3f} ms/step, t_roof fraction {big['step_roofline']['frac']}")

    cpu = None
    if args.cpu_budget > 0:
        log("CPU baseline (FP32 oracle) ...")
        picks = []
        for kind in serve.KINDS:
            r = next(q for q in reqs if q.kind == kind)
            picks.append(r.scene())
        cpu = cpu_baseline(picks, args.cpu_budget, sums)
        log(f"CPU baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} threads")

    return {
        "metric": METRIC, "value": round(sat["particle_steps_per_s"], 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": sat["batch_steps"], "warmup": args.warmup,
        "ms_per_step": round(sat["wall_s"] / max(sat["batch_steps"], 1) * 1e3, 5),
        "step_unit": "one batch step of the saturated stream (every resident scene advances one step); "
                     "steps x ms_per_step = the timed wall time",
        "steps_arg": args.steps, "timed_wall_s": round(sat["wall_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (seeded request stream, seed {args.seed}: blocks with seeded jitter / smooth velocity)",
        "config": {"workload": "served_stream", "requests": len(reqs), "rate_per_s": args.rate,
                   "mix": "50% fluid / 30% corotational / 20% svk", "particles": "log-uniform 1024-262144",
                   "steps_per_request": steps, "hbm_budget_gb": args.hbm_budget_gb, "sums": args.sums,
                   "polar": "fp64", "graph_steps": args.graph_steps, "pools": sat.get("pools", 1), "path": "batched",
                   "tuning": tuning, "warmup_requests": max(args.warmup, 1)},
        "batched_saturated": _brief(sat),
        "prefix": {"requests": len(prefix), "batched": _brief(sat_prefix), "plain": _brief(plain_prefix),
                   "speedup": round(sat_prefix["particle_steps_per_s"] / plain_prefix["particle_steps_per_s"], 3)},
        "served": {"batched": _brief(live), "plain_prefix": _brief(live_plain) if live_plain else None},
        "error_vs_single_scene": "bitwise identical (tests/test_gpu_batch.py)",
        "roofline": roof, "passes": passes, "dam_break_4m": big, "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (single-scene workloads, default 50); the served stream is fixed and "
                         "reports batch steps")
    ap.add_argument("--warmup", type=int, default=5,
                    help="warm-up steps (single scene) or warm-up requests of another seed (served stream)")
    ap.add_argument("--request-steps", type=int, default=200, help="served stream: steps per request")
    ap.add_argument("--dam-break-4m", type=int, default=1,
                    help="served stream: also measure the 4.19M-particle dam break on the plain path")
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
