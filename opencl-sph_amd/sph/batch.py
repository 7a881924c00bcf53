"""Many scenes resident on the device and stepped together.

A Batch owns a device pool (particle slots, grid cells, scene entries) and
admits scenes.Scene objects into it.  Each admitted scene advances one step
per batch step until its own step count is used up; its fields can be read
back in particle order at any time.  Results are identical to running the
scene on its own (scenes.build + run) with the same precision settings: the
batch runs the same kernels, in the same per-scene order, over each scene's
own slots and cells.

There is no counterpart in the reference, which steps one psdata at a time
(scripts/sph/systems.py:148-189); the psdata a scene is admitted from is
built from the scene's specification exactly as init_data_cpu builds it.
"""
import numpy as np

from . import bindings as bnd
from . import free_data_cpu, init_data_cpu
from .systems import field_view


def fill_fields(psdata, values):
    """Copy arrays into the leading slices of psdata fields (scalars
    broadcast), as ParticleSystem.set_fields does."""
    for name, value in values.items():
        index = bnd.get_field_psdata(psdata, name.encode())
        if index == -1:
            raise bnd.SphError(f"{name} is not a field")
        field = field_view(psdata, index)
        if isinstance(value, np.ndarray) and value.size > 1:
            np.copyto(field[tuple(slice(None, s) for s in value.shape)], value)
        else:
            field[...] = value


class Batch:
    """Device pool of `slots` particle slots, `cells` grid cells and
    `scenes` scene entries."""

    def __init__(self, slots, cells, scenes=1024, compensated_sums=0, polar_fp64=True, graph_steps=0):
        self._batch = bnd.SphBatch(slots, cells, scenes, compensated_sums, polar_fp64)
        if graph_steps:
            self._batch.set_graph_steps(graph_steps)
        self._resident = {}
        self.slots, self.cells, self.scenes = slots, cells, scenes

    @staticmethod
    def device_bytes(slots, cells, scenes=1024):
        return bnd.batch_device_bytes(slots, cells, scenes)

    @staticmethod
    def footprint(scene):
        """(slots, cells) a scene occupies in the pool."""
        gb = np.asarray(scene.params["gridbounds"], np.float32)
        h = float(np.float32(scene.params["h"]))
        # as the psdata builder: float extent, double division, truncation
        res = [int(abs(float(gb[2 * a + 1] - gb[2 * a])) / h) for a in range(3)]
        return scene.n, res[0] * res[1] * res[2] + 1

    def can_admit(self, scene):
        """Whether `scene` would be admitted now."""
        return self._batch.can_admit(*self.footprint(scene))

    @staticmethod
    def prepare(scene):
        """Host side of an admission: the scene's psdata with its initial
        fields.  Touches no device state, so it may run on another thread
        while the device steps; pass the result to add_prepared (or
        discard_prepared)."""
        psdata = init_data_cpu(scene.spec)
        try:
            fill_fields(psdata, scene.fields)
        except Exception:
            free_data_cpu(psdata)
            raise
        return psdata, scene

    @staticmethod
    def discard_prepared(prepared):
        free_data_cpu(prepared[0])

    def add_prepared(self, prepared, steps=None):
        """Admit a prepared scene; returns its handle, or None when the pool
        has no room for it now (the prepared scene stays valid)."""
        psdata, scene = prepared
        handle = self._batch.add(psdata, scene.kernel_files, scene.steps if steps is None else steps)
        if handle is not None:
            self._resident[handle] = (psdata, scene)
        return handle

    def add(self, scene, steps=None):
        """Admit `scene` for `steps` steps (default scene.steps).  Returns its
        handle, or None when the pool has no room for it now."""
        prepared = self.prepare(scene)
        try:
            handle = self.add_prepared(prepared, steps)
        except Exception:
            self.discard_prepared(prepared)
            raise
        if handle is None:
            self.discard_prepared(prepared)
        return handle

    def step(self, num_steps=1):
        self._batch.step(num_steps)

    def synchronize(self):
        self._batch.synchronize()

    def steps_left(self, handle):
        return self._batch.steps_left(handle)

    def read(self, handle, *names):
        """{name: array} of the scene's current fields, first n particles
        (copies)."""
        psdata, scene = self._resident[handle]
        self._batch.read(handle, psdata, list(names))
        out = {}
        for name in names:
            view = field_view(psdata, bnd.get_field_psdata(psdata, name.encode()))
            out[name] = np.array(view[:scene.n])
        return out

    def remove(self, handle):
        psdata, _ = self._resident.pop(handle)
        self._batch.remove(handle)
        free_data_cpu(psdata)

    def info(self):
        return self._batch.info()

    def set_profiling(self, enabled=True):
        self._batch.set_profiling(enabled)

    def reset_kernel_stats(self):
        self._batch.reset_kernel_stats()

    def kernel_stats(self):
        return self._batch.kernel_stats()

    @property
    def resident(self):
        return list(self._resident)

    def free(self):
        for handle in list(self._resident):
            psdata, _ = self._resident.pop(handle)
            free_data_cpu(psdata)
        self._batch.free()
