"""Helpers to read a psdata built by either library into plain Python data."""
import ctypes
import hashlib
import os

from sph import bindings as bnd

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libref_psdata.so")


def load_ref():
    """The reference's own build_psdata/config code (oracle/_ref), or None."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB)
    lib.build_psdata_from_string.argtypes = [ctypes.POINTER(bnd.psdata), ctypes.c_char_p]
    lib.build_psdata_from_string.restype = None
    lib.free_psdata.argtypes = [ctypes.POINTER(bnd.psdata)]
    lib.free_psdata.restype = None
    lib.load_config.argtypes = [ctypes.c_char_p]
    lib.load_config.restype = ctypes.c_int
    lib.get_config_section.argtypes = [ctypes.c_char_p]
    lib.get_config_section.restype = ctypes.c_char_p
    lib.unload_config.argtypes = []
    lib.unload_config.restype = None
    lib.set_log_level.argtypes = [ctypes.c_uint]
    lib.set_log_level(5)
    for name in ("psdata_names_size", "psdata_dimensions_size", "psdata_data_size"):
        f = getattr(lib, name)
        f.argtypes = [bnd.psdata]
        f.restype = ctypes.c_uint
    return lib


def describe(ps, lib):
    """Every metadata array and the blob bytes of a psdata."""
    nf = ps.num_fields
    names_size = lib.psdata_names_size(ps)
    dims_size = lib.psdata_dimensions_size(ps) // 4
    data_size = lib.psdata_data_size(ps)
    names_addr = ctypes.c_void_p.from_address(ctypes.addressof(ps) + bnd.psdata.names.offset).value
    names_raw = ctypes.string_at(names_addr, names_size) if nf else b""
    d = {
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
    return d


def build_with(lib, spec):
    ps = bnd.psdata()
    lib.build_psdata_from_string(ctypes.byref(ps), spec.encode())
    return ps


def layout_of(spec, lib):
    ps = build_with(lib, spec)
    try:
        return describe(ps, lib)
    finally:
        lib.free_psdata(ctypes.byref(ps))


def ours():
    return bnd._lib
