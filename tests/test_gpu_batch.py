"""Scene batch: every scene's result in a mixed batch is bitwise identical to
the same scene run on its own through the single-scene path (which the
other GPU tests pin against the oracle).

Scenes of the three systems and of different sizes and step counts are
admitted at different times, finish at different steps, are removed, and
their segments are reused by later scenes; the same sequence also runs with
the step sequence captured into hipGraphs."""
import numpy as np
import pytest

from gpu_util import compare, engine, say

pytestmark = pytest.mark.gpu

from sph import scenes  # noqa: E402

FIELDS = {"fluid": ("position", "velocity", "density", "force"),
          "corotational": ("position", "velocity", "density", "force", "stress", "rotation", "density0"),
          "svk": ("position", "velocity", "density", "force", "stress", "deformation", "density0")}


def mix_scenes():
    return {
        "fA": scenes.dam_break(12, 12, 8, 0.02, 0.04, 1e-4, seed=1, steps=10, name="fA"),
        "fB": scenes.dam_break(8, 6, 5, 0.02, 0.04, 1e-4, seed=2, gridbounds=(0, 0.5, 0, 0.5, 0, 0.5), steps=14,
                               eos_tait=False, mass_consistent=False, name="fB"),
        "cA": scenes.cantilever(12, 4, 4, steps=9, name="cA"),
        "sA": scenes.svk_cube(nside=6, steps=7, name="sA"),
        "cB": scenes.cantilever(16, 5, 5, steps=5, name="cB"),
    }



def alone(scene, compensated, polar_fp64=True):
    from gpu_util import download
    s = engine(scene, compensated=compensated, polar_fp64=polar_fp64)
    s.run(scene.steps)
    out = download(s, *FIELDS[scene.system])
    s.free()
    return {k: v[:scene.n] for k, v in out.items()}


def check_same(name, got, want):
    say(f"{name}:")
    for field, w in want.items():
        mism = compare(field, got[field], w)
        assert mism == 0, f"{name}.{field}: {mism} entries differ from the scene run alone"


def run_mix(batch, sc):
    """Staggered admissions, removals and segment reuse; returns results."""
    out = {}
    h = {"fA": batch.add(sc["fA"]), "cA": batch.add(sc["cA"])}
    batch.step(3)
    h["sA"] = batch.add(sc["sA"])
    h["fB"] = batch.add(sc["fB"])
    batch.step(8)
    for k in ("fA", "cA", "sA"):
        assert batch.steps_left(h[k]) == 0, k
        out[k] = batch.read(h[k], *FIELDS[sc[k].system])
    assert batch.steps_left(h["fB"]) == 14 - 8
    batch.remove(h["cA"])
    batch.remove(h["sA"])
    h["cB"] = batch.add(sc["cB"])
    batch.step(6)
    for k in ("fB", "cB"):
        assert batch.steps_left(h[k]) == 0, k
        out[k] = batch.read(h[k], *FIELDS[sc[k].system])
    for k in ("fA", "fB", "cB"):
        batch.remove(h[k])
    assert batch.info()["scenes"] == 0 and batch.info()["used_slots"] == 0
    return out


@pytest.mark.parametrize("graph_steps,compensated", [(0, 0), (4, 0), (4, 3)])
def test_batch_matches_scenes_alone(graph_steps, compensated):
    from sph.batch import Batch
    sc = mix_scenes()
    say(f"batch parity: graph_steps={graph_steps} compensated={compensated}")
    want = {k: alone(s, compensated) for k, s in sc.items()}
    batch = Batch(slots=1 << 14, cells=1 << 16, scenes=16, compensated_sums=compensated, graph_steps=graph_steps)
    try:
        got = run_mix(batch, sc)
        info = batch.info()
        say(f"  graph captures {info['graph_captures']}, launches {info['graph_launches']}")
        if graph_steps:
            assert info["graph_launches"] >= 1
    finally:
        batch.free()
    for k in sc:
        check_same(k, got[k], want[k])


@pytest.mark.parametrize("compensated", [0, 3])
def test_batch_dense_packing_matches_alone(compensated):
    """Scenes packed at h = 2.35 dx overflow most interior neighbour lists,
    and a crowded cell overflows the list entries' row offsets: the batched
    passes take the candidate-scan fallback on the same lanes as the
    single-scene passes, with the same results, bit for bit."""
    from sph.batch import Batch
    sc = {"fD": scenes.dam_break(12, 12, 8, 0.04 / 2.35, 0.04, 2e-3, seed=3, steps=6, name="fD"),
          "cD": scenes.cantilever(12, 8, 8, h=0.0235, steps=5, name="cD"),
          "sD": scenes.svk_cube(nside=10, h=0.0235, steps=4, name="sD"),
          # 4913 particles in one cell: rows longer than a list entry addresses
          "fC": scenes.dam_break(17, 17, 17, 0.04 / 17, 0.04, 1e-4, seed=5, origin=(0.481, 0.481, 0.481), steps=3,
                                 name="fC")}
    say(f"batch dense packing: compensated={compensated}")
    want = {k: alone(s, compensated) for k, s in sc.items()}
    batch = Batch(slots=1 << 14, cells=1 << 16, scenes=16, compensated_sums=compensated, graph_steps=4)
    try:
        h = {k: batch.add(s) for k, s in sc.items()}
        batch.step(6)
        got = {}
        for k, s in sc.items():
            assert batch.steps_left(h[k]) == 0, k
            got[k] = batch.read(h[k], *FIELDS[s.system])
            batch.remove(h[k])
    finally:
        batch.free()
    for k in sc:
        check_same(k, got[k], want[k])


def test_batch_fluid_static_body_matches_alone_and_oracle():
    """A fluid scene whose lowest layer is a static body (body_number 1,
    body_dynamic [0]): the body numbers move with the particles through
    every reorder, so the same particles stay frozen in the batch as alone
    (kernels/common.cl:699-704) and as in the FP32 oracle."""
    from gpu_util import oracle_state
    from sph.batch import Batch
    sc = {"fS": scenes.dam_break(10, 9, 7, 0.02, 0.04, 1e-4, seed=4, steps=8, static_floor=True, name="fS"),
          "cA": scenes.cantilever(12, 4, 4, steps=8, name="cA")}
    say("batch fluid with a static floor")
    want = {k: alone(s, 1) for k, s in sc.items()}
    st = oracle_state(sc["fS"], params={"compensated": 1})
    st.step("fluid", sc["fS"].steps)
    floor = sc["fS"].fields["body_number"] == 1
    assert np.array_equal(st.f["position"][floor], sc["fS"].fields["position"][floor])
    assert compare("fS position alone vs oracle", want["fS"]["position"], st.f["position"]) == 0
    batch = Batch(slots=1 << 14, cells=1 << 16, scenes=8, compensated_sums=1, graph_steps=4)
    try:
        h = {k: batch.add(s) for k, s in sc.items()}
        batch.step(8)
        got = {k: batch.read(h[k], *FIELDS[s.system]) for k, s in sc.items()}
    finally:
        batch.free()
    for k in sc:
        check_same(k, got[k], want[k])


def test_batch_admission_when_full():
    """A scene that does not fit is refused (None) and fits after removals."""
    from sph.batch import Batch
    sc = mix_scenes()
    slots = sc["fA"].n + sc["sA"].n          # cA (192) fits beside fA, sA (216) only once cA is gone
    batch = Batch(slots=slots, cells=1 << 16, scenes=4)
    try:
        a = batch.add(sc["fA"])
        c = batch.add(sc["cA"])
        assert a is not None and c is not None
        assert batch.add(sc["sA"]) is None
        batch.step(2)
        batch.remove(c)
        s = batch.add(sc["sA"])
        assert s is not None
        batch.step(sc["sA"].steps)
        got = batch.read(s, *FIELDS["svk"])
    finally:
        batch.free()
    check_same("sA after reuse", got, alone(sc["sA"], 0))


def test_batch_rejects_bad_scenes():
    from sph import bindings as bnd
    from sph.batch import Batch
    sc = mix_scenes()
    batch = Batch(slots=1 << 12, cells=1 << 14, scenes=2)
    try:
        big = scenes.dam_break(20, 20, 12, 0.02, 0.04, 1e-4, steps=1)   # 4800 particles > 4096 slots
        with pytest.raises(bnd.SphError):
            batch.add(big)
        bad = scenes.Scene("x", "fluid", sc["fA"].spec, "fluids", sc["fA"].fields, sc["fA"].params, 1)
        with pytest.raises(bnd.SphError):
            batch.add(bad)                                             # no 'common'
    finally:
        batch.free()


def test_served_stream_same_alone_and_in_mix():
    """Requests served by the continuous-batching loop (arrivals while others
    run, graph replay) give the same bits as the same requests served one at
    a time through the single-scene path."""
    from sph import serve
    reqs = serve.make_requests(count=12, rate=400.0, seed=3, nmin=300, nmax=3000, steps=12)
    say("served stream: " + ", ".join(f"{r.kind}:{r.n}" for r in reqs))
    got, want = {}, {}
    fields = ("position", "velocity")
    rep = serve.serve_batched(reqs, slots=1 << 16, cells=1 << 18, scene_entries=16, graph_steps=4, realtime=True,
                              compensated_sums=1, fields=fields, on_result=lambda r, res: got.__setitem__(r.index, res))
    say(f"  batched: {rep['batch_steps']} batch steps, peak {rep['peak_resident_scenes']} scenes resident, "
        f"graph launches {rep['graph_launches']}")
    serve.serve_plain(reqs, realtime=False, compensated_sums=1, fields=fields,
                      on_result=lambda r, res: want.__setitem__(r.index, res))
    assert sorted(got) == sorted(want) == list(range(len(reqs)))
    for r in reqs:
        check_same(f"request {r.index} ({r.kind}, {r.n})", got[r.index], want[r.index])


def test_served_results_equal_oracle():
    """Requests through the continuous-batching loop (batched steps, graph
    replay, the benchmarks' Kahan density sums and FP64 polar) against the
    FP32 oracle stepping each request's scene with the same settings: bit for
    bit."""
    from gpu_util import oracle_state
    from sph import serve
    # seed 41: 3 fluid, 4 corotational and 3 SVK requests
    reqs = serve.make_requests(count=10, rate=400.0, seed=41, nmin=300, nmax=2500, steps=10)
    say("served vs oracle: " + ", ".join(f"{r.kind}:{r.n}" for r in reqs))
    got = {}
    fields = ("position", "velocity")
    serve.serve_batched(reqs, slots=1 << 16, cells=1 << 18, scene_entries=16, graph_steps=4, realtime=False,
                        compensated_sums=1, fields=fields, on_result=lambda r, res: got.__setitem__(r.index, res))
    assert sorted(got) == list(range(len(reqs)))
    for r in reqs:
        scene = r.scene()
        st = oracle_state(scene, params={"compensated": 1})
        st.step(scene.system, r.steps)
        say(f"request {r.index} ({r.kind}, {r.n}):")
        for k in fields:
            assert compare(k, got[r.index][k], st.f[k][:scene.n]) == 0
