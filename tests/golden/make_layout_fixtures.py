"""Write tests/golden/psdata_layouts.json: the psdata layout (metadata arrays
and blob bytes) that the reference's own builder (oracle/_ref, compiled from
src/build_psdata.c, src/particle_system.c, src/stringly.c, src/config.c)
produces for every spec exercised by tests/test_psdata_layout.py, for the
$psdata_specification section of the reference's four shipped confs
("conf:reference/<file>", read from /root/reference/conf) and of this
package's confs ("conf:project/<file>").  Conf entries also hold the SHA-256
of the section text they were built from.

    make -C oracle ref && python tests/golden/make_layout_fixtures.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
REPO = os.path.dirname(TESTS)
sys.path[:0] = [TESTS, os.path.join(REPO, "opencl-sph_amd")]

import layout_util as lu  # noqa: E402
from test_psdata_layout import ALL_SPECS  # noqa: E402

REFERENCE = "/root/reference"


def main():
    ref = lu.load_ref()
    if ref is None:
        sys.exit("oracle/_ref/libref_psdata.so is missing: run `make -C oracle ref`")
    out = {name: lu.layout_of(spec, ref) for name, spec in sorted(ALL_SPECS.items())}
    for kind, folder in (("reference", os.path.join(REFERENCE, "conf")),
                         ("project", os.path.join(REPO, "opencl-sph_amd", "conf"))):
        for name in sorted(os.listdir(folder)):
            if not name.endswith(".conf"):
                continue
            assert ref.load_config(os.path.join(folder, name).encode()) == 0
            text = ref.get_config_section(b"psdata_specification")
            ref.unload_config()
            entry = lu.layout_of(text.decode(), ref)
            entry["section_sha256"] = hashlib.sha256(text).hexdigest()
            out[f"conf:{kind}/{name}"] = entry
    path = os.path.join(HERE, "psdata_layouts.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(path, len(out), "layouts")


if __name__ == "__main__":
    main()
