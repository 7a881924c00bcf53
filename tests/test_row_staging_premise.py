"""The density scans stage each neighbour row of a wave in LDS as one run of
consecutive slots: from the first scanning lane's row start to the last
scanning lane's row end (kernels_fluid.hip, k_density_listed and
k_density_solid_listed).  That is the union of the lanes' ranges only if,
over the consecutive slots of a wave, a row's start and end never decrease:
the slots are sorted by cell key, a row's first and last cells are monotone
functions of the key (x clamped to the grid, rows outside the grid skipped),
and cell offsets are an exclusive scan of the counts.  This restates that
range computation (row_range) in numpy on binned scenes, edges and
out-of-grid rows included, and checks the premise for every wave and row,
as well as the size of the staged run that the scans' stage capacity must
hold for the benchmark scenes."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "opencl-sph_amd"))
from sph import scenes  # noqa: E402


def binned(pos, gridbounds, h):
    """Slots sorted by cell key (stable), as the bin passes leave them
    (kernels/common.cl:372-391 for the key), with the grid's counts and
    exclusive-scan offsets."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    gb = np.asarray(gridbounds, np.float32).reshape(3, 2)
    h = np.float32(h)
    res = np.array([int(np.float32(abs(gb[a, 1] - gb[a, 0])) / h) for a in range(3)], np.int64)
    lo, hi = gb[:, 0], gb[:, 1]
    gc = np.clip(((pos - lo) / (hi - lo) * res).astype(np.int64), 0, res - 1)
    key = gc[:, 2] * res[0] * res[1] + gc[:, 1] * res[0] + gc[:, 0]
    key = key[np.argsort(key, kind="stable")]
    count = np.bincount(key, minlength=int(res.prod()))
    offset = np.concatenate(([0], np.cumsum(count)[:-1]))
    return key, res, count, offset


def row_ranges(key, res, count, offset):
    """For every slot and neighbour row r = 0..8 ((dz, dy) = (r // 3 - 1,
    r % 3 - 1)): whether the row lies in the grid, and its slot range
    [b, e) over cells x - 1 .. x + 1 clamped to the grid."""
    rx, ry, rz = (int(v) for v in res)
    xo, yo, zo = key % rx, (key // rx) % ry, key // (rx * ry)
    inside = np.zeros((len(key), 9), bool)
    b = np.zeros((len(key), 9), np.int64)
    e = np.zeros((len(key), 9), np.int64)
    for r in range(9):
        zn, yn = zo + r // 3 - 1, yo + r % 3 - 1
        ok = (zn >= 0) & (zn < rz) & (yn >= 0) & (yn < ry)
        row = np.where(ok, zn * rx * ry + yn * rx, 0)
        x0 = np.maximum(xo - 1, 0)
        x1 = np.minimum(xo + 1, rx - 1)
        inside[:, r] = ok
        b[:, r] = offset[row + x0]
        e[:, r] = offset[row + x1] + count[row + x1]
    return inside, b, e


def check_waves(inside, b, e, wave=64):
    """The premise, per wave and row: over the lanes whose row is inside the
    grid, starts and ends never decrease, so each lane's range lies in
    [first lane's start, last lane's end).  Returns the largest such run."""
    longest = 0
    for w0 in range(0, len(b), wave):
        for r in range(9):
            m = inside[w0:w0 + wave, r]
            if not m.any():
                continue
            bs, es = b[w0:w0 + wave, r][m], e[w0:w0 + wave, r][m]
            assert np.all(np.diff(bs) >= 0) and np.all(np.diff(es) >= 0)
            assert np.all(bs >= bs[0]) and np.all(es <= es[-1]) and np.all(bs <= es)
            longest = max(longest, int(es[-1] - bs[0]))
    return longest


def test_premise_on_a_jittered_block_with_grid_edges():
    """A block filling its grid to the edges (rows outside the grid at every
    face), particles jittered across cell boundaries."""
    rng = np.random.default_rng(3)
    pts, _ = scenes.lattice(23, 17, 11, 0.02)
    pos = pts + rng.uniform(-0.004, 0.004, pts.shape)
    gb = [-0.01, 0.45, -0.01, 0.33, -0.01, 0.21]
    check_waves(*row_ranges(*binned(pos, gb, 0.04)))


def test_premise_with_sparse_and_crowded_cells():
    """Particles scattered at random with a crowded cell and cells left
    empty: waves then span several grid rows and layers."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.0, 1.0, (6000, 3))
    pos[:700] = 0.5 + rng.uniform(-0.01, 0.01, (700, 3))
    check_waves(*row_ranges(*binned(pos, [0, 1, 0, 1, 0, 1], 0.05)))


def test_staged_runs_of_the_benchmark_scenes():
    """The staged run of a wave's row over the served stream's kinds of
    scene and the 16 k fluid: at least four in five within the stage (96
    slots fluid, 128 solid); longer runs are scanned from global memory."""
    for sc, cap in ((scenes.request_scene("fluid", 30, 20, 15, 2), 96),
                    (scenes.request_scene("corotational", 30, 25, 20, 1), 128),
                    (scenes.baseline_config(0), 96)):
        key, res, count, offset = binned(sc.fields["position"], sc.params["gridbounds"], sc.params["h"])
        inside, b, e = row_ranges(key, res, count, offset)
        longest = check_waves(inside, b, e)
        runs = []
        for w0 in range(0, len(b), 64):
            for r in range(9):
                m = inside[w0:w0 + 64, r]
                if m.any():
                    runs.append(e[w0:w0 + 64, r][m][-1] - b[w0:w0 + 64, r][m][0])
        runs = np.array(runs)
        assert (runs <= cap).mean() > 0.8, (sc.name, (runs <= cap).mean(), longest)
