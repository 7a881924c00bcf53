"""Rigid-body tests: a soft corotational solid held by scripted rigid bodies
whose positions and rotations follow per-step animations, with a per-step
history of chosen fields.

Counterpart of scripts/rigidtest.py:4-110 (RigidTest) and of the RigidBody
container of scripts/sandbox2.py:16-31, driving a StrainSolid of this
package: the same parameters, body numbering, animation semantics, history
layout and failure detection.

    left = RigidBody(particles=left_block, pos_animation=path)   # path: (steps, 3)
    test = RigidTest(soft_particles=middle_block, rigid_bodies=[left, RigidBody(particles=right_block)],
                     total_mass=10.0, strain_solid=StrainSolid(), history_particles="all",
                     history_fields=["position"], run_step_count=900)
    test.run()                      # "success" or "failure"

Particles are numbered soft particles first, then each body's particles in
order; body i (0-based) carries body_number i + 1.  Every step, body j is
moved to pos_animation[step] and turned to angle_animation[step] (Euler
angles) when those animations are given and long enough, then the solid
takes one step and the history fields are recorded.

Where the reference's file does not run as written, the evident intent is
followed:
- it calls strain_solid.setup(**fields), which its own systems.py does not
  accept; here the fields go to init(), then to the device, then setup()
  computes the reference configuration;
- it starts body i + 1 at len(soft) + len(body i), which is the cumulative
  start for two bodies; here the starts are cumulative for any number;
- history_particles other than "all" refer to an undefined name; here they
  are rows of particle positions, matched exactly as in the strain tests.
"""
import argparse

import numpy as np

from .straintest import column_intersection_indices

REQUIRED = ("soft_particles", "rigid_bodies", "total_mass", "strain_solid", "history_particles",
            "history_fields", "run_step_count")


class RigidBody(object):
    """One rigid body: its particles, its reference position and Euler angles,
    and optional per-step animations of both (scripts/sandbox2.py:16-31)."""

    def __init__(self, **kwargs):
        if "particles" not in kwargs:
            raise Exception("You need to supply particle positions")
        self.particles = kwargs["particles"]
        self.position = kwargs["position"] if "position" in kwargs else np.zeros(3)
        self.euler_angles = kwargs["euler_angles"] if "euler_angles" in kwargs else np.zeros(3)
        if "pos_animation" in kwargs:
            self.pos_animation = kwargs["pos_animation"]
        if "angle_animation" in kwargs:
            self.angle_animation = kwargs["angle_animation"]


class RigidTest(object):
    def __init__(self, **params):
        missing = [key for key in REQUIRED if key not in params]
        if missing:
            raise Exception("Missing required fields: {}".format(missing))
        self.params = argparse.Namespace(**params)
        p = self.params

        soft = np.asarray(p.soft_particles)
        sizes = [np.asarray(body.particles).shape[0] for body in p.rigid_bodies]
        self.rigid_body_start_indices = (soft.shape[0] + np.concatenate(([0], np.cumsum(sizes)[:-1]))
                                         ).astype(np.uint32) if sizes else np.zeros(0, np.uint32)
        self.positions = np.concatenate((soft,) + tuple(np.asarray(body.particles) for body in p.rigid_bodies))
        n = self.positions.shape[0]
        rotations = np.tile(np.eye(3), (n, 1, 1))
        self.body_numbers = np.zeros(n, dtype=np.uint32)
        for i, (start, size) in enumerate(zip(self.rigid_body_start_indices, sizes)):
            self.body_numbers[start:start + size] = i + 1

        fields = dict(position=self.positions, mass=p.total_mass / n, n=n, rotation=rotations,
                      body_number=self.body_numbers, gravity=np.array([0, 0, -0]))
        if p.rigid_bodies:
            fields["body_position"] = np.array([body.position for body in p.rigid_bodies])
            fields["body_euler_rotation"] = np.array([body.euler_angles for body in p.rigid_bodies])
        p.strain_solid.init(**fields)

        self._init_history_indices()
        self._init_history_arrays()

        p.strain_solid.sync_host_to_device()
        p.strain_solid.setup()
        self.body_positions = p.strain_solid.get_fields("body_position")

    def _init_history_indices(self):
        if isinstance(self.params.history_particles, str) and self.params.history_particles == "all":
            self.history_particle_indices = np.arange(self.positions.shape[0])
        else:
            self.history_particle_indices = column_intersection_indices(self.positions,
                                                                        self.params.history_particles)

    def _init_history_arrays(self):
        p = self.params
        prototypes = p.strain_solid.get_fields(*p.history_fields)
        self.history = {name: np.zeros((p.run_step_count, self.history_particle_indices.size) + proto.shape[1:])
                        for name, proto in zip(p.history_fields, prototypes)}
        self.history_counter = 0

    def detect_failure(self, **data):
        return np.any(np.isnan(data["position"]))

    def record_history(self):
        if self.history_counter >= self.params.run_step_count:
            raise Exception("History is full")
        for key in self.history:
            np.copyto(self.history[key][self.history_counter],
                      self.params.strain_solid.get_fields(key)[0][self.history_particle_indices])
        self.history_counter += 1

    def run(self):
        ps = self.params.strain_solid
        for i in range(self.params.run_step_count):
            for j, body in enumerate(self.params.rigid_bodies):
                if hasattr(body, "pos_animation") and i < body.pos_animation.shape[0]:
                    ps.set_body_position(j, body.pos_animation[i])
                if hasattr(body, "angle_animation") and i < body.angle_animation.shape[0]:
                    ps.set_body_rotation(j, body.angle_animation[i])
            ps.update()
            ps.sync_fields_device_to_host(*self.params.history_fields)
            self.record_history()
            if self.detect_failure(position=self.history["position"][self.history_counter - 1]):
                ps.sync_fields_device_to_host()
                print("Something exploded or got set to nan: stopping run")
                print("The fields from the last frame are available in params['strain_solid']")
                return "failure"
        return "success"
