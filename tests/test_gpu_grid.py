"""Neighbour search on the GPU: bit-exact cell keys, counts, offsets and
stable cell order (ascending particle index inside a cell) against the
sequential oracle, and neighbour sets against brute force."""
import numpy as np
import pytest

from gpu_util import download, engine, oracle_state, say
from sph import scenes

pytestmark = pytest.mark.gpu

GRID = ("gridcell", "gridcount", "celloffset", "cellparticles")


def check_bins(system, st, n):
    got = download(system, *GRID)
    st.bins()
    assert np.array_equal(got["gridcell"][:n], st.gridcell)
    assert np.array_equal(got["gridcount"].reshape(-1), st.gridcount)
    assert np.array_equal(got["celloffset"].reshape(-1), st.celloffset)
    assert np.array_equal(got["cellparticles"][:n], st.cellparticles)
    return got


def test_bins_dam_break_bit_exact(hip_device):
    scene = scenes.baseline_config(0)
    say(f"bins on {scene.name}, n={scene.n}")
    s = engine(scene)
    st = oracle_state(scene)
    s.compute_particle_bins()
    got = check_bins(s, st, scene.n)
    # invariants of the stable counting sort
    cp = got["cellparticles"][:scene.n]
    assert np.array_equal(np.sort(cp), np.arange(scene.n))
    assert int(got["gridcount"].sum()) == scene.n
    key = got["gridcell"][cp]
    assert np.all(np.diff(key.astype(np.int64)) >= 0)
    same_cell = np.diff(key.astype(np.int64)) == 0
    assert np.all(np.diff(cp.astype(np.int64))[same_cell] > 0)
    s.free()


def test_bins_repeated_after_motion(hip_device):
    """Bins on an already cell-sorted device state give the same result as
    a fresh sequential binning of the moved positions."""
    scene = scenes.dam_break(20, 18, 9, 0.02, 0.04, 1e-3, seed=3, gridbounds=(0, 1, 0, 1, 0, 1), steps=3)
    s = engine(scene)
    s.compute_particle_bins()
    rng = np.random.Generator(np.random.PCG64(7))
    pos = scene.fields["position"].astype(np.float64)
    for rep in range(3):
        pos = pos + rng.normal(scale=0.02, size=pos.shape)
        p32 = pos.astype(np.float32)
        s.set_fields(position=p32)
        s.compute_particle_bins()
        st = oracle_state(scene, position=p32)
        check_bins(s, st, scene.n)
        say(f"  repeat {rep} ok")
    s.free()


def test_bins_out_of_bounds_partial_and_ragged(hip_device):
    """Clamping of positions outside the grid (and NaN), n < pnum, and one
    crowded cell."""
    rng = np.random.Generator(np.random.PCG64(11))
    pnum, n = 3000, 2500
    pos = rng.uniform(-0.5, 1.5, size=(pnum, 3)).astype(np.float32)
    pos[:400] = np.float32(0.51)            # 400 particles in one cell
    pos[400, 0] = np.nan
    pos[401] = np.float32(np.inf)
    spec = scenes.fluid_spec(pnum, 0.008, 1e-4, 0.1, (0, 1, 0, 1, 0, 1), 1000, 2000, 0.1)
    scene = scenes.Scene("ragged", "fluid", spec, scenes.FLUID_KERNELS,
                         {"position": pos, "n": np.array([n], np.uint32)},
                         dict(n=n, gridbounds=[0, 1, 0, 1, 0, 1], h=0.1, mass=0.008, dt=1e-4))
    s = engine(scene)
    s.compute_particle_bins()
    st = oracle_state(scene, position=pos[:n])
    got = check_bins(s, st, n)
    assert got["gridcount"].max() >= 400
    # inactive particles keep their data
    back = download(s, "position")["position"]
    assert np.array_equal(back[n:], pos[n:])
    assert np.array_equal(back[:n], pos[:n], equal_nan=True)
    s.free()


@pytest.mark.parametrize("h", [0.0093, 0.0038], ids=["300_tiles", "4442_tiles"])
def test_bins_many_scan_tiles(hip_device, h):
    """Cell scans of 1.23 M and 18.2 M cells (scan tiles of 4096 cells, the
    last one partial): below and above the tile count (4096) up to which each
    tile sums the tiles before it itself rather than after a separate scan of
    the tile sums."""
    rng = np.random.Generator(np.random.PCG64(13))
    pnum = 6000
    pos = rng.uniform(0.0, 1.0, size=(pnum, 3)).astype(np.float32)
    pos[:300] = np.float32(0.2)              # one crowded cell
    spec = scenes.fluid_spec(pnum, 0.008, 1e-4, h, (0, 1, 0, 1, 0, 1), 1000, 2000, 0.1)
    scene = scenes.Scene("many_tiles", "fluid", spec, scenes.FLUID_KERNELS,
                         {"position": pos, "n": np.array([pnum], np.uint32)},
                         dict(n=pnum, gridbounds=[0, 1, 0, 1, 0, 1], h=h, mass=0.008, dt=1e-4))
    res = int(np.float32(1.0) / np.float32(h))
    say(f"bins on a {res}^3 grid ({res ** 3} cells), n={pnum}")
    s = engine(scene)
    s.compute_particle_bins()
    st = oracle_state(scene, position=pos)
    got = check_bins(s, st, pnum)
    assert got["gridcount"].size == res ** 3
    assert got["gridcount"].size % 4096 != 0
    s.free()


def test_bins_empty(hip_device):
    spec = scenes.fluid_spec(64, 0.008, 1e-4, 0.1, (0, 1, 0, 1, 0, 1), 1000, 2000, 0.1)
    scene = scenes.Scene("empty", "fluid", spec, scenes.FLUID_KERNELS,
                         {"position": np.full((64, 3), 0.5, np.float32), "n": np.array([0], np.uint32)},
                         dict(n=0, gridbounds=[0, 1, 0, 1, 0, 1], h=0.1, mass=0.008, dt=1e-4))
    s = engine(scene)
    s.compute_particle_bins()
    got = download(s, *GRID)
    assert int(got["gridcount"].sum()) == 0
    assert not got["celloffset"].any()
    s.run(2)
    s.free()


def test_neighbour_counts_match_brute_force(hip_device):
    """num_neighbours counts r < h over the 27-cell candidates; the cell edge
    is >= h, so it must equal the O(N^2) count with the same float32 r."""
    rng = np.random.Generator(np.random.PCG64(5))
    n = 2048
    pos = rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32)
    h = 0.09
    spec = scenes.fluid_spec(n, 0.008, 1e-4, h, (0, 1, 0, 1, 0, 1), 1000, 2000, 0.1)
    scene = scenes.Scene("nbr", "fluid", spec, scenes.FLUID_KERNELS,
                         {"position": pos, "n": np.array([n], np.uint32)},
                         dict(n=n, gridbounds=[0, 1, 0, 1, 0, 1], h=h, mass=0.008, dt=1e-4))
    s = engine(scene)
    s.compute_particle_bins()
    s.call_for_all_particles("compute_density")
    got = download(s, "num_neighbours")["num_neighbours"]
    d = pos[None, :, :] - pos[:, None, :]
    r = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    want = (r < np.float32(h)).sum(axis=1)
    assert np.array_equal(got, want)
    s.free()
