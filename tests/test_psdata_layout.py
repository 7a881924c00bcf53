"""psdata builder and .conf loader: byte-identical to the reference.

Three pins:
  * the reference's own build_psdata.c / config.c compiled from its sources
    (oracle/_ref, built by `make -C oracle ref` when the reference tree is
    present) on the shipped confs, the engine's scene specs and edge cases;
  * committed fixtures (tests/golden/psdata_layouts.json) written from that
    library by tests/golden/make_layout_fixtures.py, so the comparison also
    runs where the reference tree is absent;
  * closed-form expectations for the alignment rule.
"""
import json
import os

import numpy as np
import pytest

import layout_util as lu
from conftest import REFERENCE, have_reference
from sph import bindings as bnd
from sph import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psdata_layouts.json")

EDGE_SPECS = {
    # vec3 field first: the first offset goes through (size_t)0 - 1 and lands on 8
    "first_field_tuplet_12": """position | 3, 4 | float
pnum | 1 | unsigned int | 4
n | 1 | unsigned int
smoothingradius | 1 | float | 0.5
gridbounds | 2, 3 | float | 0, 1, 0, 2, 0, 3
""",
    # comments, blank lines, separator runs, aliases grouped last, named and hex/octal dims
    "formatting_and_aliases": """# a comment line
pnum  ,, | 1 |   unsigned   int | 6

n | 1 | unsigned int
count | 1 | uint | 3
mass | 1 | float32 | 0.25
big | 1 | float64 | 2.5
weights | 0x3, pnum | float
table | 010 | int | -1, 2, -3, 4, 5, 6, 7, 8
flags | pnum | char
ids | pnum | int
smoothingradius | 1 | float | 0.1
gridbounds | 2, 3 | float | -1, 1, -1, 1, -1, 1
velocity | 3, pnum | float
""",
    # double grid parameters, dims through a float-typed field's bits
    "double_grid": """pnum | 1 | unsigned int | 5
n | 1 | unsigned int | 5
smoothingradius | 1 | double | 0.3
gridbounds | 2, 3 | double | 0, 1.0, -0.5, 0.5, 0, 0.9
k | 1 | unsigned int | 2
pair | k, pnum | double
position | 3, pnum | float
extra | 1 | float | 1, 2, 3
""",
    # no value for smoothingradius: the grid arrays are not added
    "no_grid": """pnum | 1 | unsigned int | 3
n | 1 | unsigned int
smoothingradius | 1 | float
gridbounds | 2, 3 | float | 0, 1, 0, 1, 0, 1
position | 3, pnum | float
""",
    # more than five dimensions are cut at five
    "too_many_dims": """pnum | 1 | unsigned int | 2
n | 1 | unsigned int
smoothingradius | 1 | float | 1
gridbounds | 2, 3 | float | 0, 3, 0, 2, 0, 1
t | 1, 2, 1, 2, 1, 2, 7 | float
position | 3, pnum | float
""",
}


def scene_specs():
    return {
        "fluid_small": scenes.fluid_spec(1000, 0.008, 1e-4, 0.04, (0, 1, 0, 1, 0, 1), 1000, 2000, 0.1,
                                         eos_tait=True, mass_consistent=True),
        "corotational_small": scenes.solid_spec(512, 0.001, 5e-5, 0.02, (-0.1, 0.2, -0.1, 0.3, -0.1, 0.4),
                                                6e-5, 1e5, 0.3),
        "svk_small": scenes.solid_spec(729, 0.0011, 2e-5, 0.02, (-0.1, 0.2, -0.1, 0.2, -0.1, 0.2), 4e-4, 5e5,
                                       0.35, svk=True),
    }


ALL_SPECS = {**EDGE_SPECS, **scene_specs()}


def shipped_conf_sections():
    out = {}
    if not have_reference():
        return out
    conf_dir = os.path.join(REFERENCE, "conf")
    for name in sorted(os.listdir(conf_dir)):
        if name.endswith(".conf"):
            out[name] = os.path.join(conf_dir, name)
    return out


@pytest.fixture(scope="module")
def ref():
    lib = lu.load_ref()
    if lib is None:
        pytest.skip("oracle/_ref/libref_psdata.so not built (needs the reference sources)")
    return lib


@pytest.mark.parametrize("name", sorted(ALL_SPECS))
def test_layout_matches_reference_build(ref, name):
    spec = ALL_SPECS[name]
    assert lu.layout_of(spec, lu.ours()) == lu.layout_of(spec, ref)


@pytest.mark.parametrize("name", sorted(shipped_conf_sections()) or ["none"])
def test_shipped_confs_load_and_layout_match(ref, name, tmp_path):
    confs = shipped_conf_sections()
    if name not in confs:
        pytest.skip("reference tree not present")
    path = confs[name].encode()
    sections = {}
    for lib in (lu.ours(), ref):
        assert lib.load_config(path) == 0
        spec = lib.get_config_section(b"psdata_specification")
        kern = lib.get_config_section(b"opencl_kernel_files")
        missing = lib.get_config_section(b"no_such_section")
        sections[id(lib)] = (spec, kern, missing)
        lib.unload_config()
    (a, b) = sections.values()
    assert a == b
    assert a[2] is None
    assert lu.layout_of(a[0].decode(), lu.ours()) == lu.layout_of(a[0].decode(), ref)


CONF_TEXTS = {
    "leading_dollar": "$one\nalpha | 1 | float | 1\n$two\nx, y\n",
    "preamble": "# header\n$psdata_specification\npnum | 1 | unsigned int | 3\n$opencl_kernel_files\ncommon\n",
    "no_sections": "just text\nno dollar\n",
    "empty_body": "$a\n$b\nbody\n",
    "nine_sections": "".join(f"$s{k}\nbody{k}\n" for k in range(9)),
    "title_spaces": "$ spaced title \ncontent\n",
}


@pytest.mark.parametrize("name", sorted(CONF_TEXTS))
def test_config_sections_match_reference(ref, name, tmp_path):
    path = tmp_path / f"{name}.conf"
    path.write_text(CONF_TEXTS[name])
    titles = ["one", "two", "a", "b", "psdata_specification", "opencl_kernel_files", " spaced title ", "spaced title"]
    titles += [f"s{k}" for k in range(9)]
    results = []
    for lib in (lu.ours(), ref):
        assert lib.load_config(str(path).encode()) == 0
        results.append([lib.get_config_section(t.encode()) for t in titles])
        lib.unload_config()
    assert results[0] == results[1]


def test_missing_conf_returns_one(ref):
    assert lu.ours().load_config(b"/nonexistent/file.conf") == 1
    assert ref.load_config(b"/nonexistent/file.conf") == 1


def test_layout_fixtures_hold():
    """Committed fixture written from the reference build."""
    with open(GOLDEN) as f:
        golden = json.load(f)
    specs = {k: v for k, v in golden.items() if not k.startswith("conf:")}   # conf entries: test_psdata_tool.py
    assert set(specs) == set(ALL_SPECS)
    for name, expected in specs.items():
        assert lu.layout_of(ALL_SPECS[name], lu.ours()) == expected, name


def test_alignment_rule_closed_form():
    """offset_f = tuplet * ceil(end / tuplet), tuplet = entry * dims[0] for
    multi-dimensional fields; the first field sits at (2^64 - 1) // t * t + t
    mod 2^64."""
    lay = lu.layout_of(EDGE_SPECS["first_field_tuplet_12"], lu.ours())
    names = lay["names"].split("\0")[:-1]
    assert names[0] == "position" and lay["data_offsets"][0] == ((2**64 - 1) // 12 + 1) * 12 % 2**64 == 8
    end = 0
    for f in range(lay["num_fields"]):
        nd = lay["num_dimensions"][f]
        dims = lay["dimensions"][lay["dimensions_offsets"][f]:lay["dimensions_offsets"][f] + nd]
        t = lay["entry_sizes"][f] * (dims[0] if nd > 1 else 1)
        expect = (t * (((end - 1) % 2**64) // t + 1)) % 2**64
        assert lay["data_offsets"][f] == expect % 2**32
        end = lay["data_offsets"][f] + lay["data_sizes"][f]


def test_grid_arrays_and_type_grouping():
    lay = lu.layout_of(ALL_SPECS["fluid_small"], lu.ours())
    names = lay["names"].split("\0")[:-1]
    # grid arrays appended, then regrouped: floats first, then unsigned ints
    for g in ("gridres", "gridcell", "gridcount", "celloffset", "cellparticles"):
        assert g in names
    types = lay["types"]
    first_uint = types.index(bnd.T_UINT)
    assert all(t == bnd.T_FLOAT32 for t in types[:first_uint])
    assert all(t == bnd.T_UINT for t in types[first_uint:])
    res_index = names.index("gridres")
    ps = lu.build_with(lu.ours(), ALL_SPECS["fluid_small"])
    try:
        res = np.ctypeslib.as_array(ps.data_ptrs[res_index].u, shape=(3,)).copy()
    finally:
        bnd.free_psdata(ps)
    assert list(res) == [25, 25, 25]   # |1 - 0| / 0.04 in double, truncated
