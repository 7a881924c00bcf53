"""Python interface of the HIP SPH engine.

Same call surface as the reference's `sph` package (scripts/sph/__init__.py:
5-46): init_opencl, terminate_opencl, init_data(conf) -> (psdata, device
handle), free_data, init_data_cpu, init_data_opencl, ...  The device handle
is a HIP handle (PsdataHip); the *_opencl names are kept so that callers of
the reference run unchanged, with *_hip aliases.
"""
import ctypes

from . import bindings as bnd
from .bindings import PsdataHip, SphError

__all__ = ["init_opencl", "init_hip", "terminate_opencl", "terminate_hip", "init_data", "init_data_from_string",
           "free_data", "init_data_cpu", "free_data_cpu", "init_data_opencl", "free_data_opencl", "PsdataHip",
           "SphError"]


def init_opencl():
    """Select the HIP device (scripts/sph/__init__.py:5-6)."""
    bnd.init_hip()


def terminate_opencl():
    bnd.terminate_hip()


init_hip = init_opencl
terminate_hip = terminate_opencl


def init_data(config_file):
    """Build psdata from a .conf file and mirror it on the device
    (scripts/sph/__init__.py:11-22).  Returns None when the file cannot be
    read, as the reference does."""
    path = config_file.encode() if isinstance(config_file, str) else config_file
    if bnd.load_config(path) > 0:
        print("Initialization failed: could not read config file {}".format(config_file))
        return None
    try:
        spec = bnd.get_config_section("psdata_specification")
        kernels = bnd.get_config_section("opencl_kernel_files")
        if spec is None:
            raise SphError(f"{config_file} has no $psdata_specification section")
        psdata = init_data_cpu(spec)
        handle = init_data_opencl(psdata, kernels or "")
    finally:
        bnd.unload_config()
    return psdata, handle


def init_data_from_string(psdata_specification, kernel_files):
    """init_data from the two section texts instead of a file."""
    psdata = init_data_cpu(psdata_specification)
    return psdata, init_data_opencl(psdata, kernel_files)


def free_data(psdata, handle):
    free_data_cpu(psdata)
    free_data_opencl(handle)


def init_data_cpu(config_spec_string):
    psdata = bnd.psdata()
    spec = config_spec_string.encode() if isinstance(config_spec_string, str) else config_spec_string
    bnd.build_psdata_from_string(ctypes.byref(psdata), spec)
    if bnd.psdata_last_build_status() != 0:
        raise SphError("psdata specification could not be built (see the log above)")
    return psdata


def free_data_cpu(psdata):
    bnd.free_psdata(ctypes.byref(psdata))


def init_data_opencl(psdata, config_kernels):
    return PsdataHip(psdata, config_kernels)


def free_data_opencl(handle):
    handle.free()


init_data_hip = init_data_opencl
free_data_hip = free_data_opencl
