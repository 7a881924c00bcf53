"""Autotuner: launch parameters per device and workload, chosen by timing.

Search space:

    neighbour_block   threads per block of the neighbour-list passes and
                      candidate scans: 64, 128, 256
    graph_steps       batch steps per hipGraph (batches only): 2, 4, 8, 16
    pair_group        list entries per load group of the fluid force pass:
                      8, 12, 16 (tuned after the other two, with their winners)

The cell size is not a tuning parameter: it is the reference's grid
(resolution |hi - lo| / h per axis), which fixes the neighbour visiting
order and therefore the results.  Nothing tuned here changes a result bit;
only timings differ.

Method: every candidate runs a short representative workload once untimed,
then `repeats` timed runs; the median decides.  Winners are kept in a JSON
table keyed by device (architecture and compute units, "gfx950 256CU", from
sph_hip_device_name) and workload signature:

    single:<system>:<particles>      one scene, stepped through the plain path
    batch:<mix>                      the served stream's request mix

Lookup order: $SPH_HIP_TUNING (a JSON file), ~/.cache/sph_hip/tuning.json,
then the table shipped with the package (sph/tuning.json); tune() writes to
the first of the first two.
"""
import json
import os
import statistics
import time

from . import bindings as bnd

HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGED = os.path.join(HERE, "tuning.json")
BLOCKS = (64, 128, 256)
GRAPH_STEPS = (2, 4, 8, 16)
PAIR_GROUPS = (8, 12, 16)
DEFAULTS = {"neighbour_block": 256, "graph_steps": 8, "pair_group": 8}


def user_table_path():
    return os.environ.get("SPH_HIP_TUNING") or os.path.join(os.path.expanduser("~"), ".cache", "sph_hip",
                                                            "tuning.json")


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(signature, device=None):
    """Tuned parameters for `signature` on this device (defaults when none)."""
    device = device or bnd.device_name()
    for path in (user_table_path(), PACKAGED):
        entry = _read(path).get(device, {}).get(signature)
        if entry:
            return dict(DEFAULTS, **entry["params"])
    return dict(DEFAULTS)


def apply(params):
    bnd.set_tuning(params["neighbour_block"], params.get("pair_group", DEFAULTS["pair_group"]))


def _store(signature, params, timings, device):
    path = user_table_path()
    table = _read(path)
    table.setdefault(device, {})[signature] = {"params": params, "timings_s": timings,
                                               "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S")}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def _median_time(run, repeats):
    """Median over `repeats` of run()'s own measure (seconds) when it returns
    one, else of its wall time; one untimed run first."""
    run()                                    # untimed: first-launch effects
    times = []
    for _ in range(repeats):
        t = time.perf_counter()
        measured = run()
        times.append(measured if measured is not None else time.perf_counter() - t)
    return statistics.median(times)


def single_signature(scene):
    return f"single:{scene.system}:{scene.n}"


def tune_single(scene, steps=10, repeats=3, blocks=BLOCKS, pair_groups=PAIR_GROUPS, store=True, log=None):
    """Neighbour block size (and, for fluids, the force pass's load group)
    for one scene on the single-scene path."""
    from . import scenes
    system = scenes.build(scene)
    handle = system.psdata_opencl
    timings = {}

    def run():
        system.run(steps)
        handle.synchronize()
    try:
        for nb in blocks:
            bnd.set_tuning(nb, DEFAULTS["pair_group"])
            timings[str(nb)] = _median_time(run, repeats)
            if log:
                log(f"tune {single_signature(scene)}: block {nb}: {timings[str(nb)] / steps * 1e3:.3f} ms/step")
        best_nb = int(min(timings, key=timings.get))
        best_pg = DEFAULTS["pair_group"]
        if scene.system == "fluid":
            group_t = {DEFAULTS["pair_group"]: timings[str(best_nb)]}
            for pg in pair_groups:
                if pg in group_t:
                    continue
                bnd.set_tuning(best_nb, pg)
                group_t[pg] = _median_time(run, repeats)
                timings[f"{best_nb}/g{pg}"] = group_t[pg]
                if log:
                    log(f"tune {single_signature(scene)}: block {best_nb}, pair group {pg}: "
                        f"{group_t[pg] / steps * 1e3:.3f} ms/step")
            best_pg = min(group_t, key=group_t.get)
    finally:
        system.free()
    params = dict(DEFAULTS, neighbour_block=best_nb, pair_group=int(best_pg))
    apply(params)
    if store:
        _store(single_signature(scene), params, timings, bnd.device_name())
    return params, timings


def batch_signature(mix="50/30/20"):
    return f"batch:{mix}"


def tune_batch(requests, blocks=BLOCKS, graph_steps=GRAPH_STEPS, pair_groups=PAIR_GROUPS, repeats=2,
               compensated_sums=1, store=True, signature=None, log=None):
    """Neighbour block size and steps per graph for a batch serving
    `requests` saturated (all queued at once)."""
    from . import serve
    timings = {}
    # a pool sized to the sample: the time measured is the serving loop's
    # own wall time, from the first admission to the last result
    total = sum(r.n for r in requests)
    slots = max(1 << 16, 2 * total)
    def runner(k):
        def run():
            return serve.serve_batched(requests, realtime=False, graph_steps=k, slots=slots, cells=8 * slots,
                                       compensated_sums=compensated_sums)["wall_s"]
        return run
    for nb in blocks:
        bnd.set_tuning(nb, DEFAULTS["pair_group"])
        for k in graph_steps:
            key = f"{nb}/{k}"
            timings[key] = _median_time(runner(k), repeats)
            if log:
                log(f"tune batch: block {nb}, graph steps {k}: {timings[key]:.3f} s")
    best = min(timings, key=timings.get)
    nb, k = (int(x) for x in best.split("/"))
    group_t = {DEFAULTS["pair_group"]: timings[best]}
    for pg in pair_groups:
        if pg in group_t:
            continue
        bnd.set_tuning(nb, pg)
        group_t[pg] = _median_time(runner(k), repeats)
        timings[f"{nb}/{k}/g{pg}"] = group_t[pg]
        if log:
            log(f"tune batch: block {nb}, graph steps {k}, pair group {pg}: {group_t[pg]:.3f} s")
    params = dict(DEFAULTS, neighbour_block=nb, graph_steps=k, pair_group=int(min(group_t, key=group_t.get)))
    apply(params)
    if store:
        _store(signature or batch_signature(), params, timings, bnd.device_name())
    return params, timings
