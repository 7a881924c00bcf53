"""Served scene stream: requests arrive over time, are admitted into a device
batch as slots and memory free up, run their steps together with whatever
else is resident, and are read back as soon as they finish.

    requests = make_requests(count=2048, rate=40.0, seed=0)
    report = serve_batched(requests, hbm_budget=200e9)
    report = serve_plain(requests)          # one scene at a time, for comparison

A request is a seeded scene description (system, block size, seed, steps);
its scene is generated when it is admitted.  Latency is the time from a
request's arrival to its results being in host memory.  With realtime=False
every request is available at time 0 (a saturated queue), which measures
throughput alone.

There is no counterpart in the reference (which steps one psdata at a time
from Python, scripts/sph/systems.py:148-189); the per-scene step sequence is
the reference's, through the batch (sph.batch) or the single-scene path
(scenes.build).
"""
import concurrent.futures
import dataclasses
import time
from collections import deque

import numpy as np

from . import scenes

KINDS = ("fluid", "corotational", "svk")


@dataclasses.dataclass
class Request:
    index: int
    arrival: float          # seconds after the start of the stream
    kind: str
    dims: tuple
    seed: int
    steps: int

    @property
    def n(self):
        return int(np.prod(self.dims))

    def scene(self):
        return scenes.request_scene(self.kind, *self.dims, seed=self.seed, steps=self.steps,
                                    name=f"request_{self.index}")


def make_requests(count=2048, rate=40.0, seed=0, mix=(0.5, 0.3, 0.2), nmin=1024, nmax=262144, steps=200):
    """Seeded request stream: exponential inter-arrival times at `rate` per
    second, system drawn from `mix`, particle count log-uniform in
    [nmin, nmax] as a block of random aspect."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        kind = KINDS[int(rng.choice(3, p=np.asarray(mix) / np.sum(mix)))]
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        dims = scenes.block_dims(n, rng)
        out.append(Request(i, t, kind, dims, int(rng.integers(0, 2**31 - 1)), steps))
    return out


def _summary(name, requests, done, t_start, t_end, extra):
    lat = np.array([done[r.index]["latency"] for r in requests])
    work = float(sum(r.n * r.steps for r in requests))
    by_system = {k: float(sum(r.n * r.steps for r in requests if r.kind == k)) for k in KINDS}
    wall = t_end - t_start
    rep = {
        "path": name, "requests": len(requests), "particle_steps": work, "particle_steps_by_system": by_system,
        "wall_s": wall,
        "particle_steps_per_s": work / wall if wall > 0 else 0.0,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_mean_s": float(lat.mean()), "latency_max_s": float(lat.max()),
    }
    rep.update(extra)
    return rep


class _Preparer:
    """Scene generation and psdata building on worker threads, for requests
    that have arrived, at most `ahead` prepared and not yet admitted.  The
    device is only touched by the serving thread."""

    def __init__(self, threads, ahead):
        from .batch import Batch
        self._batch_cls = Batch
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=threads)
        self.futures = {}
        self.ahead = ahead

    def submit(self, req):
        if req.index not in self.futures and len(self.futures) < self.ahead:
            self.futures[req.index] = self.pool.submit(lambda r=req: self._batch_cls.prepare(r.scene()))

    def close(self):
        for fut in self.futures.values():
            if not fut.cancel():
                try:
                    self._batch_cls.discard_prepared(fut.result())
                except Exception:
                    pass
        self.pool.shutdown(wait=True)
        self.futures.clear()


def serve_batched(requests, hbm_budget=200e9, slots=None, cells=None, scene_entries=4096, graph_steps=8,
                  compensated_sums=1, polar_fp64=True, realtime=True, on_result=None, fields=("position",),
                  prepare_threads=2, prepare_ahead=64, profile=False, pools=2, log=None):
    """Continuous batching over device pools.

    Each pool is sized to its share of the HBM budget (or to `slots` /
    `cells` when given, within the budget).  Worker threads generate the
    scenes of arrived requests and build their psdata while the device
    steps.  Each round of the loop admits into one pool every prepared
    request that fits (in arrival order, looking past ones that do not fit
    yet) and enqueues `graph_steps` batch steps for it, then waits for the
    round enqueued before on the next pool and reads back and releases the
    scenes that finished there.  With two pools (the default) the second
    takes requests while the first is full: under a saturated queue the
    device then runs one pool's round while the host serves the other and
    does not wait for the host between rounds, and a light load runs in one
    pool as with `pools=1`.  Profiling uses one pool and waits for every
    preparation it reaches, so that a profiled pass admits the same scenes
    in the same rounds however fast the workers run.
    on_result(request, {field: array}) receives each result; `fields` names
    the fields read back, for every request or per system ({kind: names})."""
    from .batch import Batch

    npool = 1 if profile else max(1, int(pools))
    slots = slots or 32 << 20
    cells = cells or 2 * slots
    while npool * Batch.device_bytes(slots, cells, scene_entries) > hbm_budget and slots > 1024:
        slots //= 2
        cells //= 2
    pool_bytes = npool * Batch.device_bytes(slots, cells, scene_entries)
    batches = [Batch(slots, cells, scene_entries, compensated_sums=compensated_sums, polar_fp64=polar_fp64,
                     graph_steps=0 if profile else graph_steps) for _ in range(npool)]
    prep = _Preparer(prepare_threads, prepare_ahead)
    if profile:
        batches[0].set_profiling(True)
    stats = {}
    chunk = max(graph_steps, 1)
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    waiting = []                          # arrived, not admitted, in arrival order
    running = [{} for _ in range(npool)]  # per pool: handle -> request
    in_flight = [False] * npool           # a round enqueued and not yet collected
    done = {}
    busy_steps = 0
    peak_slots = peak_scenes = 0
    phase = {"prepare_wait": 0.0, "admit": 0.0, "enqueue": 0.0, "wait": 0.0, "read": 0.0, "idle": 0.0}
    t0 = time.perf_counter()
    clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))

    def arrivals():
        now = clock()
        while pending and pending[0].arrival <= now:
            waiting.append(pending.popleft())
        for req in waiting:
            prep.submit(req)

    full = [False] * npool                # the pool's last admission met a request that did not fit

    def admit(k):
        """Admission into pool k: arrival order, looking past requests that do
        not fit yet; stops at the first request still being prepared.  A pool
        after the first admits only while the one before it is full, so a
        light load stays in one pool (one pool's round is the shortest)."""
        nonlocal waiting
        if k > 0 and not full[k - 1]:
            return
        batch = batches[k]
        admitted = set()
        misses = 0
        for req in waiting:
            fut = prep.futures.get(req.index)
            if fut is None:
                break
            if not fut.done():
                if (any(running) or admitted) and not profile:
                    break
                tp = time.perf_counter()
                concurrent.futures.wait([fut])       # nothing to step: wait for the scene
                phase["prepare_wait"] += time.perf_counter() - tp
            prepared = fut.result()
            if not batch.can_admit(prepared[1]):
                if not running[k] and not admitted:
                    from .bindings import SphError
                    raise SphError(f"request {req.index} does not fit the empty pool "
                                   f"({slots} slots, {cells} cells)")
                misses += 1
                if misses >= 4:        # the pool is close to full
                    break
                continue
            ta = time.perf_counter()
            handle = batch.add_prepared(prepared, req.steps)
            phase["admit"] += time.perf_counter() - ta
            if handle is None:
                continue
            del prep.futures[req.index]
            running[k][handle] = req
            admitted.add(req.index)
        full[k] = misses > 0
        if admitted:
            waiting = [r for r in waiting if r.index not in admitted]

    def collect(k):
        """Wait for pool k's round; read back and release its finished scenes."""
        if not in_flight[k]:
            return
        batch = batches[k]
        tw = time.perf_counter()
        batch.synchronize()
        phase["wait"] += time.perf_counter() - tw
        in_flight[k] = False
        tr = time.perf_counter()
        for handle in [h for h in running[k] if batch.steps_left(h) == 0]:
            req = running[k].pop(handle)
            result = batch.read(handle, *(fields[req.kind] if isinstance(fields, dict) else fields))
            batch.remove(handle)
            t_done = time.perf_counter() - t0
            done[req.index] = {"latency": t_done - (req.arrival if realtime else 0.0), "finished": t_done}
            if on_result:
                on_result(req, result)
        phase["read"] += time.perf_counter() - tr

    try:
        k = 0
        while pending or waiting or any(running):
            arrivals()
            admit(k)
            if running[k]:
                infos = [b.info() for b in batches]
                peak_slots = max(peak_slots, sum(i["used_slots"] for i in infos))
                peak_scenes = max(peak_scenes, sum(i["scenes"] for i in infos))
                te = time.perf_counter()
                batches[k].step(chunk)
                in_flight[k] = True
                busy_steps += chunk
                phase["enqueue"] += time.perf_counter() - te
                arrivals()                     # preparation continues on the workers
            nxt = (k + 1) % npool
            collect(nxt)
            if not any(running) and not any(in_flight):
                if pending and not waiting:
                    ti = time.perf_counter()
                    time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.05)))
                    phase["idle"] += time.perf_counter() - ti
            k = nxt
            if log and busy_steps and busy_steps % (chunk * 100) == 0:
                log(f"batched: {len(done)}/{len(requests)} done, {sum(len(r) for r in running)} resident, "
                    f"{len(waiting)} waiting, t={time.perf_counter() - t0:.1f}s")
        for j in range(npool):
            collect(j)
        t_end = time.perf_counter() - t0
        infos = [b.info() for b in batches]
        if profile:
            stats = batches[0].kernel_stats()
    finally:
        prep.close()
        for b in batches:
            b.free()
    t_first = requests[0].arrival if realtime else 0.0
    wall = max(t_end - min(t_first, t_end), 1e-9)
    return _summary("batched", requests, done, min(t_first, t_end), t_end,
                    {# the host blocked on the device for `wait`, so the device was executing
                     # at least that share of the wall time; prepare_wait: the host waited for
                     # scene preparation (psdata builds) with the device possibly idle
                     "gpu_busy_lower_bound": round(phase["wait"] / wall, 4),
                     "prepare_wait_share": round(phase["prepare_wait"] / wall, 4),
                     "pool_bytes": pool_bytes, "pool_slots": slots, "pool_cells": cells, "pools": npool,
                     "peak_resident_slots": peak_slots, "peak_resident_scenes": peak_scenes,
                     "batch_steps": busy_steps, "graph_captures": sum(i["graph_captures"] for i in infos),
                     "graph_launches": sum(i["graph_launches"] for i in infos), "graph_steps": graph_steps,
                     "kernel_stats": stats, "host_phases_s": {k: round(v, 4) for k, v in phase.items()}})


def serve_plain(requests, compensated_sums=1, polar_fp64=True, realtime=True, on_result=None, fields=("position",),
                log=None):
    """One scene at a time through the single-scene path (its own psdata and
    device handle, set-up, steps, readback, release), in arrival order."""
    done = {}
    t0 = time.perf_counter()
    for k, req in enumerate(sorted(requests, key=lambda r: r.arrival)):
        if realtime:
            wait = req.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        scene = req.scene()
        system = scenes.build(scene, setup=False)
        system.psdata_opencl.set_precision(compensated_sums, polar_fp64)
        if scene.system in ("corotational", "svk"):
            system.setup()
        system.run(req.steps)
        system.sync_fields_device_to_host(*fields)
        result = {f: np.array(system.get_fields(f)[0][:scene.n]) for f in fields}
        system.free()
        t_done = time.perf_counter() - t0
        done[req.index] = {"latency": t_done - (req.arrival if realtime else 0.0), "finished": t_done}
        if on_result:
            on_result(req, result)
        if log and (k + 1) % 100 == 0:
            log(f"plain: {k + 1}/{len(requests)} done, t={t_done:.1f}s")
    t_end = time.perf_counter() - t0
    t_first = requests[0].arrival if realtime else 0.0
    return _summary("plain", requests, done, min(t_first, t_end), t_end, {})
