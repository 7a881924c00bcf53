"""Benchmark: particle-steps/s of the SPH step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dam_break_4m]

A step is one full simulation step (neighbour grid, density, forces,
integrator) over the whole scene, with the state resident in HBM.  Prints one
JSON line with the throughput, the roofline of the dominant kernel (HIP-event
timing of each kernel over a second pass of the same K steps) and a CPU
baseline (the FP32 oracle on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "opencl-sph_amd"))

METRIC = "particle-steps/s batched vs plain at fixed FP64-oracle error; p50/p99 latency"
WORKLOADS = {"dam_break_16k": 0, "cantilever_16k": 1, "svk_cube_262k": 2, "dam_break_4m": 3}


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


# Timed kernel name -> device symbol prefix in the rocprofv3 counter summaries.
KERNEL_SYMBOLS = {"compute_density": "k_density_listed", "compute_forces_fluids": "k_forces_fluids_listed",
                  "bin_rank_place": "k_rank_place", "step_forward": "k_step_forward",
                  "bin_scatter": "k_scatter_unsorted", "bin_key_count": "k_key_count"}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the calibrated FETCH_SIZE /
    WRITE_SIZE summary committed under profiles/, or None."""
    path = os.path.join(REPO, "profiles", f"{workload}_hbm_traffic.json")
    prefix = KERNEL_SYMBOLS.get(kernel)
    if prefix is None or not os.path.exists(path):
        return None
    with open(path) as f:
        summary = json.load(f)
    for name, row in summary.get("kernels", {}).items():
        if name.startswith(prefix):
            return int(row["read_bytes"] + row["write_bytes"])
    return None


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


def cpu_baseline(scene, budget_s):
    """The FP32 oracle on the host cores, full scene, as many whole steps as
    fit the time budget (at least one after a warm-up step)."""
    import ctypes

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    fields = {k: v for k, v in scene.fields.items() if k in oracle.OracleState.PER_PARTICLE}
    st = oracle.OracleState("f32", scene.params, **fields)
    lib, _, _ = oracle.load("f32")
    lib.orc_threads.restype = ctypes.c_uint
    threads = int(lib.orc_threads())
    st.step(scene.system, 1)                        # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        st.step(scene.system, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= scene.steps:
            break
    return {"value": scene.n * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} step(s) of the full {scene.n}-particle {scene.name} scene, FP32 oracle "
                      f"(C restatement of the reference kernels, OpenMP), after 1 warm-up step, "
                      f"{el:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dam_break_4m", choices=sorted(WORKLOADS))
    ap.add_argument("--sums", default="kahan-density", choices=["plain", "kahan-density", "kahan-all"],
                    help="neighbour-sum precision (default: Kahan-compensated density sums)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-baseline work (0: skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a device copy of this many MiB (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the engine runs one scene set per device and this build has no "
              f"multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import sph
    from sph import bindings, roofline, scenes

    bindings.set_log_level(3)          # library notes would share stdout with the JSON line
    sph.init_hip()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
    scene = scenes.baseline_config(WORKLOADS[args.workload])
    log(f"workload {scene.name}: {scene.n} particles, system {scene.system}")
    system = scenes.build(scene)
    handle = system.psdata_opencl
    handle.set_precision(compensated_sums={"plain": 0, "kahan-density": 1, "kahan-all": 3}[args.sums],
                         polar_fp64=True)

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
                "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg_s * 1e3, 5)}
    extra = {}
    if scene.system == "fluid":
        candidates, in_range = grid_pair_counts(system, scene)
        flops = roofline.flops_per_launch(dominant, candidates, in_range)
        if roof is not None and flops:
            roof["valu_tflops"] = round(flops / avg_s / 1e12, 3)
            roof["valu_frac"] = round(flops / avg_s / 1e12 / roofline.MI355X_FP32_TFLOPS, 4)
        t_roof = roofline.t_roof_seconds(scene.n, cells, candidates, in_range)
        extra = {"step_roofline": {"t_roof_ms": round(t_roof * 1e3, 5), "frac": round(t_roof * 1e3 / ms_per_step, 4),
                                   "algorithmic_bytes_per_step": roofline.step_bytes(scene.n, cells),
                                   "candidate_pairs": candidates, "in_range_pairs": in_range},
                 "kernel_ms_per_step": {k: round(v, 5) for k, v in per_step.items()}}
    system.free()

    cpu = None
    if args.cpu_budget > 0:
        log("CPU baseline (FP32 oracle) ...")
        cpu = cpu_baseline(scene, args.cpu_budget)
        log(f"CPU baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded lattice with uniform jitter)",
        "config": {"workload": scene.name, "particles": scene.n, "cells": cells, "system": scene.system,
                   "path": "plain", "sums": args.sums,
                   "fluid_model": "tait eos, mass-consistent force" if scene.system == "fluid" else None},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
