"""Standalone .conf -> psdata builder (src/build_psdata.c:40-636,
src/config.c:19-106 behind it).

    python -m sph.psdata_tool CONF [--section NAME] [--show] [--layout OUT.json]
                              [--blob OUT.bin] [--compare FIXTURES.json --key KEY]

Loads CONF with the library's config reader, builds the psdata of its
$psdata_specification section (or --section) and
  --show     prints one line per field: name, dimensions, type, entry size,
             blob offset and size (the facts display_psdata prints, without values);
  --layout   writes every metadata array, the blob's SHA-256 and the section
             text's SHA-256 as JSON, the format of the conf layout fixtures (field
             order, dimension arrays, offsets, pointer offsets, type tags);
  --blob     writes the data blob (all field values, in blob order);
  --compare  checks the layout against entry KEY of a fixture file and exits
             with status 1, listing the differing arrays, when they differ.
Needs no GPU.
"""
import argparse
import ctypes
import hashlib
import json
import sys

from . import bindings as bnd

TYPE_NAMES = {0: "char", 1: "unsigned int", 2: "int", 3: "float", 4: "double"}   # FieldType, src/datatypes.h:4-11


def _size(lib, name, ps):
    fn = getattr(lib, name)
    fn.argtypes = [bnd.psdata]
    fn.restype = ctypes.c_uint
    return fn(ps)


def describe(ps, lib=None):
    """Every metadata array of a psdata and the SHA-256 of its blob, as plain
    Python data.  `lib` is the library that built it (default: this one)."""
    lib = lib or bnd._lib
    nf = ps.num_fields
    names_size = _size(lib, "psdata_names_size", ps)
    dims_size = _size(lib, "psdata_dimensions_size", ps) // 4
    data_size = _size(lib, "psdata_data_size", ps)
    names_addr = ctypes.c_void_p.from_address(ctypes.addressof(ps) + bnd.psdata.names.offset).value
    names_raw = ctypes.string_at(names_addr, names_size) if nf else b""
    return {
        "num_fields": nf,
        "names": names_raw.decode("latin-1"),
        "names_offsets": [ps.names_offsets[i] for i in range(nf)],
        "dimensions": [ps.dimensions[i] for i in range(dims_size)],
        "num_dimensions": [ps.num_dimensions[i] for i in range(nf)],
        "dimensions_offsets": [ps.dimensions_offsets[i] for i in range(nf)],
        "entry_sizes": [ps.entry_sizes[i] for i in range(nf)],
        "data_sizes": [ps.data_sizes[i] for i in range(nf)],
        "data_offsets": [ps.data_offsets[i] for i in range(nf)],
        "types": [ps.data_ptrs[i].type for i in range(nf)],
        "ptr_offsets": [(ps.data_ptrs[i].ptr or 0) - (ps.data or 0) for i in range(nf)],
        "sizes": [names_size, dims_size * 4, data_size],
        "blob_sha256": hashlib.sha256(ctypes.string_at(ps.data, data_size) if data_size else b"").hexdigest(),
    }


def blob_bytes(ps, lib=None):
    lib = lib or bnd._lib
    size = _size(lib, "psdata_data_size", ps)
    return ctypes.string_at(ps.data, size) if size else b""


def field_table(layout):
    """(name, dims, type, entry size, offset, size) per field of a layout."""
    rows = []
    for i in range(layout["num_fields"]):
        start = layout["names_offsets"][i]
        name = layout["names"][start:layout["names"].index("\0", start)]
        d0 = layout["dimensions_offsets"][i]
        dims = layout["dimensions"][d0:d0 + layout["num_dimensions"][i]]
        rows.append((name, dims, TYPE_NAMES.get(layout["types"][i], str(layout["types"][i])),
                     layout["entry_sizes"][i], layout["data_offsets"][i], layout["data_sizes"][i]))
    return rows


def build_from_conf(path, section="psdata_specification"):
    """(psdata, section text) of a .conf file; raises bindings.SphError."""
    if bnd.load_config(path.encode()) != 0:
        raise bnd.SphError(f"cannot read config file {path}")
    try:
        text = bnd.get_config_section(section)
    finally:
        bnd.unload_config()
    if text is None:
        raise bnd.SphError(f"{path} has no ${section} section")
    ps = bnd.psdata()
    bnd.build_psdata_from_string(ctypes.byref(ps), text.encode())
    if bnd.psdata_last_build_status() != 0:
        bnd.free_psdata(ctypes.byref(ps))
        raise bnd.SphError(f"the ${section} section of {path} does not build")
    return ps, text


def differences(a, b):
    return sorted(k for k in set(a) | set(b) if a.get(k) != b.get(k))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m sph.psdata_tool", description=__doc__.splitlines()[0])
    ap.add_argument("conf")
    ap.add_argument("--section", default="psdata_specification")
    ap.add_argument("--show", action="store_true")
    ap.add_argument("--layout")
    ap.add_argument("--blob")
    ap.add_argument("--compare")
    ap.add_argument("--key")
    args = ap.parse_args(argv)
    if args.compare and not args.key:
        ap.error("--compare needs --key")
    bnd.set_log_level(2)                      # keep the library's level-1 notes off stdout
    try:
        ps, text = build_from_conf(args.conf, args.section)
    except bnd.SphError as e:
        print(f"error: {e}", file=sys.stderr)
        return 2
    try:
        layout = describe(ps)
        layout["section_sha256"] = hashlib.sha256(text.encode()).hexdigest()
        if args.show:
            for name, dims, typ, es, off, size in field_table(layout):
                print(f"{name:32s} {'x'.join(map(str, dims)) or '-':>14s} {typ:14s} {es:2d} B  "
                      f"offset {off:9d}  size {size:9d}")
            print(f"{layout['num_fields']} fields, blob {layout['sizes'][2]} bytes, sha256 {layout['blob_sha256']}")
        if args.layout:
            with open(args.layout, "w") as f:
                json.dump(layout, f, indent=1, sort_keys=True)
        if args.blob:
            with open(args.blob, "wb") as f:
                f.write(blob_bytes(ps))
        if args.compare:
            with open(args.compare) as f:
                want = json.load(f).get(args.key)
            if want is None:
                print(f"error: {args.compare} has no entry {args.key}", file=sys.stderr)
                return 2
            diff = differences(layout, want)
            if diff:
                print(f"layout differs from {args.key}: {', '.join(diff)}")
                return 1
            print(f"layout identical to {args.key}")
        return 0
    finally:
        bnd.free_psdata(ctypes.byref(ps))


if __name__ == "__main__":
    sys.exit(main())
