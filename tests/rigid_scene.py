"""A small version of the reference's rigid-body demo (scripts/sandbox2.py:
a soft block between two rigid blocks, the left one animated): lattice
spacing 0.1 m at the default smoothing radius of 0.2 m.  Test
infrastructure shared by the CPU and GPU harness tests."""
import numpy as np

from sph.rigidtest import RigidBody


def sandbox_params(strain_solid, steps=10, rotate=True):
    x, y, z = np.meshgrid(np.linspace(-0.75, 0.75, num=16), np.linspace(-0.2, 0.2, num=5),
                          np.linspace(-0.1, 0.1, num=3), indexing="ij")
    every = np.vstack((x.flatten(), y.flatten(), z.flatten())).T
    soft = every[np.logical_and(every[:, 0] > -0.4, every[:, 0] < 0.4)]
    left = every[every[:, 0] <= -0.4]
    right = every[every[:, 0] >= 0.4]
    path = np.vstack((np.linspace(0.0, -0.02, num=steps), np.zeros(steps), np.linspace(0.0, 0.01, num=steps))).T
    bodies = [RigidBody(particles=left, pos_animation=path)]
    if rotate:
        turn = np.vstack((np.zeros(steps), np.linspace(0.0, 0.05, num=steps), np.zeros(steps))).T
        bodies.append(RigidBody(particles=right, angle_animation=turn))
    else:
        bodies.append(RigidBody(particles=right))
    return {"soft_particles": soft, "rigid_bodies": bodies, "total_mass": 10.0, "strain_solid": strain_solid,
            "history_particles": "all", "history_fields": ["position", "velocity"], "run_step_count": steps}
