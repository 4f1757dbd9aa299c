"""The free rigid body of the coupling loop (CPU) against closed forms:
semi-implicit Euler translation under a constant force, a spin about a
principal axis under a constant moment, and torque-free spin about a fixed
axis (Rodrigues' rotation)."""
import math

import numpy as np

from metakonga_amd.coupling import RigidBody, rotation_matrix


def test_translation_closed_form():
    dt, n = 1e-3, 57
    g = np.array([0.0, -9.80665, 0.0])
    F = np.array([0.3, 0.1, -0.2])
    b = RigidBody(2.0, (1.0, 1.0, 1.0), origin=(1.0, 2.0, 3.0), vel=(0.5, 0.0, -0.1))
    for _ in range(n):
        b.advance(F, (0, 0, 0), g, dt)
    a = F / 2.0 + g
    assert np.allclose(b.vel, np.array([0.5, 0.0, -0.1]) + n * dt * a, rtol=1e-13, atol=1e-15)
    want = np.array([1.0, 2.0, 3.0]) + n * dt * np.array([0.5, 0.0, -0.1]) + dt * dt * a * n * (n + 1) / 2
    assert np.allclose(b.origin, want, rtol=1e-13, atol=1e-15)


def _rodrigues(axis, th):
    k = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


def test_torque_free_spin_is_an_exact_rotation():
    w = np.array([0.3, -1.2, 0.7])
    b = RigidBody(1.0, (0.2, 0.2, 0.2), origin=(0, 0, 0), omega=w)
    dt, n = 1e-2, 300
    for _ in range(n):
        b.advance((0, 0, 0), (0, 0, 0), (0, 0, 0), dt)
    assert np.allclose(b.omega, w, rtol=0, atol=1e-12)   # rounding of the frame round trip
    R = rotation_matrix(b.ep)
    assert np.abs(R - _rodrigues(w, n * dt * np.linalg.norm(w))).max() < 1e-12


def test_moment_about_a_principal_axis():
    # body turned 90 deg about x: its body z axis is the world -y axis
    ep = np.array([math.cos(math.pi / 4), math.sin(math.pi / 4), 0.0, 0.0])
    I = (0.5, 0.7, 0.9)
    b = RigidBody(1.0, I, origin=(0, 0, 0), ep=ep)
    axis = rotation_matrix(ep) @ np.array([0.0, 0.0, 1.0])
    assert np.allclose(axis, [0, -1, 0], atol=1e-15)
    M = 0.09 * axis
    dt, n = 1e-3, 400
    for _ in range(n):
        b.advance((0, 0, 0), M, (0, 0, 0), dt)
    # w stays on the principal axis (w x I w = 0): w_n = n dt M / I_z
    assert np.allclose(b.omega, n * dt * 0.09 / I[2] * axis, rtol=1e-12, atol=1e-15)
    th = sum(k * dt * 0.09 / I[2] * dt for k in range(1, n + 1))
    assert np.abs(rotation_matrix(b.ep) - _rodrigues(axis, th) @ rotation_matrix(ep)).max() < 1e-12
