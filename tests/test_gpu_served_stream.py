"""The served stream at its own scale: the first 12 requests of the seed-0
stream of 2048 (1275 to 183222 particles, all three systems) served through
the continuous-batching loop in its default layout (two pools of 32 M slots
and 64 M cells within the 200 GB budget), hipGraph replay of 8 steps, Kahan
density sums and FP64 polar decomposition, 10 steps each; every returned
field of every request is compared bit for bit with the FP32 oracle stepping
the same scene with the same settings.  The reference's per-scene step is
scripts/sph/systems.py:148-189 (kernel order of
src/opencl/particle_system_host.c:361-528)."""
import pytest

from gpu_util import compare, oracle_state, say

pytestmark = pytest.mark.gpu

FIELDS = {"fluid": ("position", "velocity", "density", "force"),
          "corotational": ("position", "velocity", "density", "force", "stress", "rotation", "strain", "density0"),
          "svk": ("position", "velocity", "density", "force", "stress", "deformation", "density0")}


def test_seed0_stream_prefix_default_layout_equals_oracle(hip_device):
    from sph import serve
    reqs = serve.make_requests(count=2048, seed=0, steps=10)[:12]
    assert max(r.n for r in reqs) > 100_000
    say("seed-0 stream prefix: " + ", ".join(f"{r.kind}:{r.n}" for r in reqs))
    got = {}
    rep = serve.serve_batched(reqs, hbm_budget=200e9, graph_steps=8, compensated_sums=1, polar_fp64=True,
                              realtime=False, fields=FIELDS,
                              on_result=lambda r, res: got.__setitem__(r.index, res))
    say(f"  pools {rep['pools']} x {rep['pool_slots']} slots / {rep['pool_cells']} cells, "
        f"{rep['batch_steps']} batch steps, graph launches {rep['graph_launches']}")
    assert rep["pool_slots"] == 32 << 20 and rep["pools"] == 2
    assert rep["graph_launches"] >= 1
    assert sorted(got) == [r.index for r in reqs]
    for r in reqs:
        scene = r.scene()
        st = oracle_state(scene, params={"compensated": 1})
        st.step(scene.system, r.steps)
        say(f"request {r.index} ({r.kind}, {r.n}):")
        for k in FIELDS[r.kind]:
            assert compare(k, got[r.index][k], st.f[k][:scene.n]) == 0, (r.index, k)
