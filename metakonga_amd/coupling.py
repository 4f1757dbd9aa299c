"""DEM–rigid-body coupling over the hand-over entry points, in the order of
the reference's solver loop (source/xDynamicsSolver.cpp:185-210): the DEM step,
then the bodies advance under the contact reactions of that step
(contact_particles_polygonObjects.cpp:469-479 hands the force to the body),
then the new poses and velocities go back to the triangles
(contactManager::update -> updatePolygonObjectData,
contact_particles_polygonObjects.cpp:258-292).

The reference couples to its multibody solver (joints, constraints), which is
out of scope here; `RigidBody` is a free body (6 DOF, principal inertia in its
own frame) so that a polygon object can move under the particles' reactions.
Translation: semi-implicit Euler, v += dt (F/m + g), x += dt v.  Rotation: the
body-frame Euler equations advanced the same way, I w' = M - w x I w, then the
orientation is turned by the exact rotation of |w| dt about w.  The reaction
moment is taken about the object's origin (its centre of mass).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def _quat_mul(a, b):
    """Hamilton product, scalar first."""
    a0, av = a[0], np.asarray(a[1:])
    b0, bv = b[0], np.asarray(b[1:])
    return np.r_[a0 * b0 - av @ bv, a0 * bv + b0 * av + np.cross(av, bv)]


def rotation_matrix(ep):
    """A(ep) of Euler parameters (e0, e1, e2, e3), the rows of makeTFM_1..3
    (mphysics_cuda_impl.cuh:624-638)."""
    e0, e1, e2, e3 = ep
    return np.array([
        [2 * (e0 * e0 + e1 * e1 - 0.5), 2 * (e1 * e2 - e0 * e3), 2 * (e1 * e3 + e0 * e2)],
        [2 * (e1 * e2 + e0 * e3), 2 * (e0 * e0 + e2 * e2 - 0.5), 2 * (e2 * e3 - e0 * e1)],
        [2 * (e1 * e3 - e0 * e2), 2 * (e2 * e3 + e0 * e1), 2 * (e0 * e0 + e3 * e3 - 0.5)]])


@dataclass
class RigidBody:
    mass: float
    inertia: tuple                       # principal moments in the body frame
    origin: np.ndarray
    ep: np.ndarray = field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0]))
    vel: np.ndarray = field(default_factory=lambda: np.zeros(3))
    omega: np.ndarray = field(default_factory=lambda: np.zeros(3))   # world frame

    def __post_init__(self):
        self.origin = np.asarray(self.origin, float).copy()
        self.ep = np.asarray(self.ep, float).copy()
        self.vel = np.asarray(self.vel, float).copy()
        self.omega = np.asarray(self.omega, float).copy()

    def advance(self, force, moment, gravity, dt):
        self.vel = self.vel + dt * (np.asarray(force, float) / self.mass + np.asarray(gravity, float))
        self.origin = self.origin + dt * self.vel
        A = rotation_matrix(self.ep)
        I = np.asarray(self.inertia, float)
        wb = A.T @ self.omega
        mb = A.T @ np.asarray(moment, float)
        wb = wb + dt * (mb - np.cross(wb, I * wb)) / I
        self.omega = A @ wb
        wn = float(np.sqrt(self.omega @ self.omega))
        if wn > 0:
            h = 0.5 * wn * dt
            q = np.r_[np.cos(h), np.sin(h) * self.omega / wn]
            self.ep = _quat_mul(q, self.ep)
            self.ep /= np.sqrt(self.ep @ self.ep)


def couple(sim, bodies: dict, nsteps: int, gravity=None):
    """Advance `sim` (metakonga_amd.xdem.Simulation) and the bodies
    {polygon object index: RigidBody} together, one step at a time."""
    g = np.asarray(sim.scene.gravity if gravity is None else gravity, float)
    dt = sim.scene.dt
    for i, b in bodies.items():
        sim.set_object_state(i, b.origin, b.ep, b.vel, b.omega)
    for _ in range(nsteps):
        sim.step(1)
        objs = sim.objects()
        for i, b in bodies.items():
            o = objs[i]
            b.advance((o.force.x, o.force.y, o.force.z), (o.moment.x, o.moment.y, o.moment.z),
                      g, dt)
            sim.set_object_state(i, b.origin, b.ep, b.vel, b.omega)
