"""global_sums_noalloc_hip (the allocation-free variant of global_sums_hip,
no counterpart in the reference) against global_sums_hip through the C ABI:
two calls in a row on caller-owned device arrays, the second with a smaller
input, each bitwise equal to global_sums_hip over a handle whose slots hold
the same particles in the same order (a fluid scene before any binning keeps
particle i in slot i)."""
import ctypes

import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory the test owns (the entry point under test allocates none)."""
    hip = None

    def __init__(self, nbytes):
        if DeviceBuffer.hip is None:
            DeviceBuffer.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(nbytes, 4))) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) == 0

    def download(self, arr):
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(arr.nbytes), 2) == 0
        return arr

    def free(self):
        self.hip.hipFree(self.ptr)


def handle_sums(scene):
    from sph import scenes
    system = scenes.build(scene)                 # uploads position and velocity, no binning
    momentum, kinetic, centre = system.psdata_opencl.global_sums()
    system.free()
    return np.array([*momentum, kinetic, *centre], np.float64)


def test_noalloc_sums_equal_handle_sums_two_calls(hip_device):
    from sph import bindings as bnd
    from sph import scenes
    fn = bnd._lib.global_sums_noalloc_hip
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    rng = np.random.Generator(np.random.PCG64(5))
    big = scenes.dam_break(20, 16, 12, 0.02, 0.04, 1e-4, seed=6, steps=1, name="big")
    small = scenes.dam_break(9, 7, 5, 0.02, 0.04, 1e-4, seed=7, steps=1, name="small")
    for sc in (big, small):
        sc.fields["velocity"] = rng.normal(size=(sc.n, 3)).astype(np.float32)
    pos = DeviceBuffer(big.n * 12)
    vel = DeviceBuffer(big.n * 12)
    out = DeviceBuffer(7 * 8)
    try:
        for sc in (big, small):                      # the second call has the smaller input
            want = handle_sums(sc)
            pos.upload(sc.fields["position"])
            vel.upload(sc.fields["velocity"])
            say(f"global_sums_noalloc_hip: {sc.name}, n = {sc.n}")
            assert fn(pos.ptr, vel.ptr, sc.n, ctypes.c_float(sc.params["mass"]), out.ptr, None) == 0, \
                bnd._last_error().decode()
            got = out.download(np.zeros(7, np.float64))
            say(f"  got {got}")
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
            m = np.float64(np.float32(sc.params["mass"]))
            v = sc.fields["velocity"].astype(np.float64)
            assert np.allclose(got[:3], m * v.sum(0), rtol=1e-12)
            assert np.allclose(got[4:], sc.fields["position"].astype(np.float64).mean(0), rtol=1e-12)
        # sizes beyond the maximum are refused before any launch
        assert fn(pos.ptr, vel.ptr, 0xFFFFFFFF, ctypes.c_float(1.0), out.ptr, None) == -1
        assert "maximum" in bnd._last_error().decode()
    finally:
        for b in (pos, vel, out):
            b.free()
