"""Strain tests: a solid block stretched and compressed between moving
constraint planes, with per-step history and per-strain output.

Counterpart of scripts/straintest.py:18-374 (StrainTest, StrainTestSym,
StrainTestSymPadding), driving a StrainSolid or SVKSolid of this package:
the same parameters, run path, plane-constraint layout, failure detection
and on-disk formats (<field>_history.npy and <field>_output.npy, numpy
memmaps, when output_folder is given).

    test = StrainTest(strain_range=np.array([0.05, 0.1, -0.05]), axis=0, strain_rate=0.5,
                      stabilisation_time=0.02, initial_dimensions=np.array([1.0, 0.5, 0.5]),
                      particle_dimensions=np.array([11, 6, 6]), total_mass=1.0,
                      strain_solid=StrainSolid(), output_particles=np.array([[10, 0, 0]]),
                      output_fields=["position"], history_particles="all",
                      history_fields=["position"])
    test.run()                      # "success" or "failure"

Run path: the constrained faces move (strain_rate x half-width per unit
time, one plane position per step) to each tension target (strain >= 0) in
the order given and hold for stabilisation_time, then the solid is reset
and the same for the compression targets in reverse order.  After every step
the plane constraints are applied and the history fields are recorded; at
the end of each hold the output fields of output_particles are recorded.
"""
import argparse
import errno
import math
import os

import numpy as np

REQUIRED = ("strain_range", "axis", "strain_rate", "stabilisation_time", "initial_dimensions",
            "particle_dimensions", "total_mass", "strain_solid", "output_particles", "output_fields",
            "history_particles", "history_fields")


def column_intersection_indices(x, y):
    """Indices of the rows of x that equal some row of y
    (scripts/straintest.py:13-16)."""
    x = np.asarray(x)
    y = np.atleast_2d(np.asarray(y))
    hit = np.zeros(len(x), bool)
    for row in y:
        hit |= np.all(x == row, axis=1)
    return np.nonzero(hit)[0]


def lattice_positions(lo, hi, counts, order="ij"):
    """Particles on a regular lattice with `counts` points per axis between lo
    and hi (inclusive).  order "ij": x slowest, z fastest; "xy": y slowest,
    then x, then z (numpy's default meshgrid order, which the reference uses
    for the padding layer).  The order is the particle numbering, which sets
    the order of the neighbour sums."""
    axes = [np.linspace(lo[a], hi[a], num=int(counts[a])) for a in range(3)]
    grid = np.meshgrid(*axes, indexing=order)
    return np.stack([g.ravel() for g in grid], axis=1)


class StrainTest(object):
    def __init__(self, **params):
        missing = [k for k in REQUIRED if k not in params]
        if missing:
            raise Exception("Missing required fields: {}".format(missing))
        self.params = argparse.Namespace(**params)
        p = self.params
        p.particle_dimensions = np.asarray(p.particle_dimensions)
        p.initial_dimensions = np.asarray(p.initial_dimensions, dtype=np.float64)
        p.strain_range = np.asarray(p.strain_range, dtype=np.float64)
        if hasattr(p, "output_folder"):
            try:
                os.makedirs(p.output_folder)
            except OSError as e:
                if not (e.errno == errno.EEXIST and os.path.isdir(p.output_folder)):
                    raise
        solid = p.strain_solid
        self.init_positions()
        self.init_particle_system()
        self.post_init()
        self.current_strain = 0
        # with the FP32 timestep field as numpy promotes it (float32 product,
        # then float64), so the plane positions are the reference's
        dt = solid.get_fields("timestep")[0]
        self.plane_frame_step = float((p.strain_rate * dt * p.initial_dimensions[p.axis] / 2)[0])
        self.stabilisation_steps = int((p.stabilisation_time / dt)[0])
        self.init_output_and_history_indices()
        self.init_run_path()
        self.init_output_arrays()
        self.init_history_arrays()
        solid.sync_host_to_device()

    # -- set-up -----------------------------------------------------------------
    def particle_mass(self):
        return self.params.total_mass / np.prod(self.params.particle_dimensions - 1)

    def init_positions(self):
        d = self.params.initial_dimensions
        self.positions = lattice_positions(-d / 2, d / 2, self.params.particle_dimensions)

    def init_particle_system(self):
        values = dict(position=self.positions, n=self.positions.shape[0], mass=self.particle_mass())
        if hasattr(self.params, "smoothingradius"):
            values["smoothingradius"] = self.params.smoothingradius
        self.params.strain_solid.init(**values)

    def post_init(self):
        self.init_plane_constraints()

    def init_plane_constraints(self):
        # plane_constraints[c, 0, :] = point on plane c, [c, 1, :] = its normal
        self.plane_constraints, self.plane_constraints_particles = \
            self.params.strain_solid.get_fields("plane_constraints", "plane_constraints_particles")
        axis = self.params.axis
        half = self.params.initial_dimensions[axis] / 2
        self.plane_constraints[0, :, axis] = [-half, 1]
        self.plane_constraints[1, :, axis] = [half, 1]
        self.plane_constraints_particles[np.nonzero(self.positions[:, axis] == -half)[0]] = 1 << 0
        self.plane_constraints_particles[np.nonzero(self.positions[:, axis] == half)[0]] = 1 << 1

    def get_init_plane_position(self):
        return self.params.initial_dimensions[self.params.axis] / 2

    def set_plane_constraints(self, plane_position):
        self.plane_constraints[0, 0, self.params.axis] = -plane_position
        self.plane_constraints[1, 0, self.params.axis] = plane_position

    def init_output_and_history_indices(self):
        counts = self.params.particle_dimensions
        grid = np.meshgrid(*[np.arange(int(c)) for c in counts], indexing="ij")
        indices = np.stack([g.ravel() for g in grid], axis=1)
        self.output_particle_indices = column_intersection_indices(indices, self.params.output_particles)
        if isinstance(self.params.history_particles, str) and self.params.history_particles == "all":
            self.history_particle_indices = np.arange(self.positions.shape[0])
        else:
            self.history_particle_indices = column_intersection_indices(indices, self.params.history_particles)

    def _stages(self, start, targets, sign):
        """One strip per target: the plane moving from the previous position
        by plane_frame_step per step, then the hold."""
        strips = []
        prev = start
        for target in targets:
            move = sign * np.arange(sign * prev, sign * target, self.plane_frame_step)
            hold = np.full(self.stabilisation_steps, target)
            strips.append(np.concatenate((move, hold)))
            prev = target
        return strips

    def init_run_path(self):
        p = self.params
        tension = np.nonzero(p.strain_range >= 0)[0]
        compression = np.nonzero(p.strain_range < 0)[0][::-1]
        self.strain_order = np.concatenate((tension, compression))
        half = self.get_init_plane_position()
        self.run_path = [self._stages(half, (p.strain_range[tension] + 1) * half, 1.0),
                         self._stages(half, (p.strain_range[compression] + 1) * half, -1.0)]
        self.run_step_count = sum(len(strip) for section in self.run_path for strip in section)

    def _arrays(self, fields, rows, indices, suffix):
        p = self.params
        protos = p.strain_solid.get_fields(*fields)
        out = {}
        for name, proto in zip(fields, protos):
            shape = (rows, len(indices)) + proto.shape[1:]
            if hasattr(p, "output_folder"):
                out[name] = np.lib.format.open_memmap(os.path.join(p.output_folder, f"{name}_{suffix}.npy"),
                                                      dtype=proto.dtype, mode="w+", shape=shape)
            else:
                out[name] = np.zeros(shape)
        return out

    def init_output_arrays(self):
        self.output = self._arrays(self.params.output_fields, self.params.strain_range.size,
                                   self.output_particle_indices, "output")

    def init_history_arrays(self):
        self.history = self._arrays(self.params.history_fields, self.run_step_count,
                                    self.history_particle_indices, "history")
        self.history_counter = 0

    # -- recording --------------------------------------------------------------
    def record_output(self, sample_index):
        for name, arr in self.output.items():
            np.copyto(arr[sample_index], self.params.strain_solid.get_fields(name)[0][self.output_particle_indices])

    def record_history(self):
        if self.history_counter >= self.run_step_count:
            raise Exception("History is full")
        for name, arr in self.history.items():
            np.copyto(arr[self.history_counter],
                      self.params.strain_solid.get_fields(name)[0][self.history_particle_indices])
        self.history_counter += 1

    def detect_failure(self, **data):
        return bool(np.any(np.isnan(data["position"])))

    # -- run ------------------------------------------------------------------------
    def run(self, **kwargs):
        p = self.params
        solid = p.strain_solid
        synced = list(dict.fromkeys(list(p.output_fields) + list(p.history_fields)))
        strip_index = 0
        solid.setup()
        for section in self.run_path:
            for strip in section:
                for plane_position in strip:
                    self.set_plane_constraints(plane_position)
                    solid.sync_fields_host_to_device("plane_constraints")
                    solid.update()
                    solid.call_for_all_particles("apply_plane_constraints")
                    solid.sync_fields_device_to_host(*synced)
                    self.record_history()
                    if "position" in self.history and \
                            self.detect_failure(position=self.history["position"][self.history_counter - 1]):
                        solid.sync_fields_device_to_host()
                        print("Something exploded or got set to nan: stopping run")
                        print("The fields from the last frame are available in params['strain_solid']")
                        return "failure"
                print("Strain: {}".format(p.strain_range[self.strain_order[strip_index]]))
                self.record_output(self.strain_order[strip_index])
                strip_index += 1
            solid.reset()
            self.post_init()
            solid.sync_host_to_device()
            solid.setup()
        return "success"


class StrainTestSym(StrainTest):
    """One octant of the block with mirror planes on x = y = z = 0
    (scripts/straintest.py:249-300)."""

    def particle_mass(self):
        return self.params.total_mass / np.prod(2 * (self.params.particle_dimensions - 1))

    def init_positions(self):
        d = self.params.initial_dimensions
        self.positions = lattice_positions(np.zeros(3), d / 2, self.params.particle_dimensions)

    def init_plane_constraints(self):
        solid = self.params.strain_solid
        self.plane_constraints, self.plane_constraints_particles, self.symmetry_planes = \
            solid.get_fields("plane_constraints", "plane_constraints_particles", "symmetry_planes")
        axis = self.params.axis
        others = ((axis + 1) % 3, (axis + 2) % 3)
        # the three faces through the origin are held on their planes
        self.plane_constraints[0, :, axis] = [0, 1]
        self.plane_constraints[1, :, others[0]] = [0, 1]
        self.plane_constraints[2, :, others[1]] = [0, 1]
        for bit, a in enumerate((axis,) + others):
            self.plane_constraints_particles[np.nonzero(self.positions[:, a] == 0)[0]] |= 1 << bit
        self.apply_moving_constraint()
        for a in range(3):
            self.symmetry_planes[a, :, a] = [0, 1]

    def apply_moving_constraint(self):
        axis = self.params.axis
        half = self.params.initial_dimensions[axis] / 2
        self.plane_constraints[3, :, axis] = [half, 1]
        self.plane_constraints_particles[np.nonzero(self.positions[:, axis] == half)[0]] |= 1 << 3

    def set_plane_constraints(self, plane_position):
        self.plane_constraints[3, 0, self.params.axis] = plane_position


class StrainTestSymPadding(StrainTestSym):
    """StrainTestSym with a layer of padding particles, at least h thick,
    beyond the free faces (scripts/straintest.py:302-374)."""

    def _spacing(self):
        p = self.params
        return [p.initial_dimensions[a] / 2.0 / (p.particle_dimensions[a] - 1) for a in range(3)]

    def init_positions(self):
        p = self.params
        d = p.initial_dimensions
        counts = p.particle_dimensions
        core = lattice_positions(np.zeros(3), d / 2, counts)
        spacing = self._spacing()
        # h / spacing in float32, as the reference computes it with the
        # numpy of its time (an FP32 field divided by a scalar stays FP32);
        # h = 0.2 over a 0.05 spacing is then exactly 4 layers
        h = np.float32(p.strain_solid.get_fields("smoothingradius")[0][0])
        layers = [int(math.ceil(h / np.float32(s))) for s in spacing]
        self.padding_width = [n * s for n, s in zip(layers, spacing)]
        full = lattice_positions(np.zeros(3), d / 2 + np.asarray(self.padding_width),
                                 [counts[a] + layers[a] for a in range(3)], order="xy")
        beyond = np.zeros(len(full), bool)
        for a in range(3):
            beyond |= full[:, a] > (d[a] + spacing[a]) / 2
        self.positions = np.concatenate((core, full[beyond]))

    def post_init(self):
        super(StrainTestSymPadding, self).post_init()
        p = self.params
        d = p.initial_dimensions
        spacing = self._spacing()
        axis = p.axis
        self.padding_field = p.strain_solid.get_fields("padding")[0]
        for bit, a in enumerate((axis, (axis + 1) % 3, (axis + 2) % 3)):
            self.padding_field[np.nonzero(self.positions[:, a] > (d[a] + spacing[a]) / 2)[0]] |= 1 << bit

    def apply_moving_constraint(self):
        axis = self.params.axis
        half = self.params.initial_dimensions[axis] / 2
        self.plane_constraints[3, :, axis] = [self.get_init_plane_position(), 1]
        self.plane_constraints_particles[np.nonzero(self.positions[:, axis] == half + self.padding_width[axis])[0]] |= 1 << 3

    def get_init_plane_position(self):
        return self.params.initial_dimensions[self.params.axis] / 2 + self.padding_width[self.params.axis]
