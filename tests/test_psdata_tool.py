"""The standalone .conf -> psdata tool (sph.psdata_tool) against the layout
fixtures that the reference's own builder produced (tests/golden): this
package's confs everywhere, the reference's four shipped confs where the
reference tree is present."""
import hashlib
import json
import os
import subprocess
import sys

import pytest

from conftest import PKG, REFERENCE, REPO, have_reference

FIXTURES = os.path.join(REPO, "tests", "golden", "psdata_layouts.json")


def run_tool(*args):
    return subprocess.run([sys.executable, "-m", "sph.psdata_tool", *args], cwd=PKG, capture_output=True,
                          text=True, timeout=120)


def conf_cases():
    out = [("project", os.path.join(PKG, "conf", n)) for n in sorted(os.listdir(os.path.join(PKG, "conf")))
           if n.endswith(".conf")]
    if have_reference():
        folder = os.path.join(REFERENCE, "conf")
        out += [("reference", os.path.join(folder, n)) for n in sorted(os.listdir(folder)) if n.endswith(".conf")]
    return out


def test_fixtures_cover_every_conf():
    keys = json.load(open(FIXTURES))
    assert {k for k in keys if k.startswith("conf:reference/")} == {
        "conf:reference/" + n for n in ("fluid.conf", "solid.conf", "strainsolid.conf", "svksolid.conf")}
    for kind, path in conf_cases():
        assert f"conf:{kind}/{os.path.basename(path)}" in keys


@pytest.mark.parametrize("kind,path", conf_cases(), ids=lambda v: os.path.basename(v) if "/" in str(v) else v)
def test_tool_layout_equals_fixture(kind, path, tmp_path):
    key = f"conf:{kind}/{os.path.basename(path)}"
    out = tmp_path / "layout.json"
    blob = tmp_path / "blob.bin"
    res = run_tool(path, "--layout", str(out), "--blob", str(blob), "--compare", FIXTURES, "--key", key)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "identical" in res.stdout
    got = json.load(open(out))
    want = json.load(open(FIXTURES))[key]
    assert got == want
    data = blob.read_bytes()
    assert len(data) == got["sizes"][2]
    assert hashlib.sha256(data).hexdigest() == got["blob_sha256"]


def test_tool_reports_differences_and_errors(tmp_path):
    conf = os.path.join(PKG, "conf", "fluid.conf")
    res = run_tool(conf, "--compare", FIXTURES, "--key", "conf:project/svksolid.conf")
    assert res.returncode == 1 and "differs" in res.stdout
    res = run_tool(conf, "--show")
    assert res.returncode == 0 and "position" in res.stdout and "fields, blob" in res.stdout
    bad = tmp_path / "bad.conf"
    bad.write_text("$other\nx\n")
    res = run_tool(str(bad))
    assert res.returncode == 2 and "no $psdata_specification" in res.stderr
    res = run_tool(str(tmp_path / "missing.conf"))
    assert res.returncode == 2
