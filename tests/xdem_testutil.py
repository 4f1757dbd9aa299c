"""Helpers shared by the tests: device buffers through the C ABI, contact-rich
states, and the cancellation-safe force error metric."""
from __future__ import annotations

import ctypes as C

import numpy as np

from metakonga_amd import abi, scenes


class DBuf:
    """A device buffer allocated through xdem_malloc (no torch)."""

    def __init__(self, lib, host: np.ndarray | None = None, nbytes: int | None = None):
        self.lib = lib
        if host is not None:
            host = np.ascontiguousarray(host)
            nbytes = host.nbytes
        self.nbytes = max(int(nbytes), 1)
        p = C.c_void_p()
        abi.check(lib, lib.xdem_malloc(C.byref(p), self.nbytes))
        self.ptr = p
        if host is not None and host.nbytes:
            abi.check(lib, lib.xdem_memcpy_htod(self.ptr, C.c_void_p(host.ctypes.data), host.nbytes))

    def get(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            abi.check(self.lib, self.lib.xdem_memcpy_dtoh(C.c_void_p(out.ctypes.data), self.ptr,
                                                           out.nbytes))
        return out

    def set(self, host):
        host = np.ascontiguousarray(host)
        abi.check(self.lib, self.lib.xdem_memcpy_htod(self.ptr, C.c_void_p(host.ctypes.data),
                                                       host.nbytes))

    def __del__(self):
        try:
            self.lib.xdem_free(self.ptr)
        except Exception:
            pass


def struct_dev(lib, ctype, items):
    arr = abi.struct_array(ctype, items)
    raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
    return DBuf(lib, raw), arr


def dense_state(n_side=12, radius=5e-3, pitch_factor=1.96, jitter=0.03, seed=11, vel_scale=0.05,
                omega_scale=5.0, f32_exact=True):
    """A compressed jittered lattice with random velocities and spins: every
    particle has several overlapping neighbours (contact-rich state)."""
    r = float(scenes.as_f32_exact(radius))
    n = n_side ** 3
    idx = np.arange(n)
    ix, iy, iz = idx % n_side, (idx // n_side) % n_side, idx // (n_side * n_side)
    pitch = pitch_factor * r
    u = scenes.uniform(seed, 9 * n, 3).reshape(n, 9) * 2 - 1
    pos = np.stack([0.05 + ix * pitch + u[:, 0] * jitter * r,
                    -0.5 + iy * pitch + u[:, 1] * jitter * r,
                    0.05 + iz * pitch + u[:, 2] * jitter * r, np.full(n, r)], 1)
    vel = u[:, 3:6] * vel_scale
    om = u[:, 6:9] * omega_scale
    if f32_exact:
        vel = scenes.as_f32_exact(vel)
        om = scenes.as_f32_exact(om)
    return np.ascontiguousarray(pos), np.ascontiguousarray(vel), np.ascontiguousarray(om)


def force_error(f_test, f_ref, f_scale):
    """max_i |f_test - f_ref|_inf / scale_i, with scale_i = sum of the magnitudes
    of the terms that were summed into particle i (the raw net force of a
    settled particle cancels)."""
    d = np.abs(np.asarray(f_test) - np.asarray(f_ref)).max(axis=1)
    return float((d / np.maximum(f_scale, 1e-300)).max()) if d.size else 0.0


def dense_scene(n_side=12, seed=11, pitch=1.97, jitter=0.04, vel_scale=0.05, omega_scale=5.0):
    """A contact-rich dense scene in an open box of planes, with its random
    velocities and spins (FP32-representable)."""
    pos, vel, om = dense_state(n_side=n_side, seed=seed, pitch_factor=pitch, jitter=jitter,
                               vel_scale=vel_scale, omega_scale=omega_scale)
    r = pos[0, 3]
    lo = pos[:, :3].min(0) - r
    hi = pos[:, :3].max(0) + r
    g = 1
    while g * 2 * r < (hi - lo).max() + 8 * r:
        g *= 2
    sc = scenes.Scene(name="dense", pos=pos, dt=1e-5, nsteps=10, world_origin=tuple(lo - 2 * r),
                      cell_size=2 * r, grid_size=(g, g, g),
                      planes=scenes.open_box_planes(lo + 0.3 * r, hi + [0, 0.05, 0]))
    return sc, vel, om


def drum_history_scene(n_side=9, seed=51, omega=3.0, margin=0.05, planes=True, **kw):
    """The dense lattice of dense_scene inside a drum spinning about z, history
    on: the drum's end caps press on the lattice's top and bottom layers and
    its shell on the particles of the four vertical edges (overlaps up to
    margin radii); with planes, the open box of planes as well (plane springs
    come before drum springs)."""
    sc, vel, om = dense_scene(n_side=n_side, seed=seed, **kw)
    pos, r = sc.pos, sc.pos[0, 3]
    c = 0.5 * (pos[:, :3].min(0) + pos[:, :3].max(0))
    d = pos[:, :3] - c
    half = np.abs(d[:, 2]).max() + (1 - margin) * r
    rad = np.hypot(d[:, 0], d[:, 1]).max() + (1 - margin) * r
    sc.drums = [scenes.drum_info(tuple(c), (0, 0, 1), rad, half, omega, True)]
    if not planes:
        sc.planes = []
    sc.tangential_history = True
    return sc, vel, om


# ------------------------------------------------ the older entry points
def setb_params(sc, np_, nsp=0, kn=0.0, vn=0.0, ks=0.0, vs=0.0, mu=0.3, cohesive=0.0):
    """Parameter block of the older entry points (include/xdem_setb.h) for a
    scene: grid and time from the scene, pair constants as given."""
    from metakonga_amd import abi
    p = abi.BDeviceParameters()
    p.np = np_
    p.nsp = nsp
    gs = sc.grid_size
    p.ncell = gs[0] * gs[1] * gs[2]
    p.grid_size = abi.uint3(*gs)
    p.dt = sc.dt
    p.half2dt = 0.5 * sc.dt * sc.dt
    p.cell_size = sc.cell_size or 2.0 * float(sc.pos[:, 3].max())
    p.gravity = abi._d3(sc.gravity)
    p.world_origin = abi._d3(sc.world_origin)
    p.kn, p.vn, p.ks, p.vs = kn, vn, ks, vs
    p.mu = mu
    p.cohesive = cohesive
    return p


def pair_constants(orc, sc):
    """DHS constants of two particles of the scene's first radius."""
    m, _ = sc.mass_inertia()
    r = float(sc.pos[0, 3])
    c = orc.force_constant(1, r, r, float(m[0]), float(m[0]), sc.contact_pp())
    return c.kn, c.vn, c.ks, c.vs


def as4(a3, w):
    """(n, 3) + per-row w -> (n, 4) double4 layout."""
    out = np.zeros((a3.shape[0], 4))
    out[:, :3] = a3
    out[:, 3] = w
    return out


# ------------------------------------------------- FP32c error against the oracle
# North-star bar of the fused FP32c path: force (and torque) error against the
# FP64 oracle at most 1e-5, per particle, relative to
#   the term scale  sum_k (|fsn| + |fca| + (|vn| + |vs|) V_k + |Ft|) + m |g|,
# the magnitudes of the terms the contact law combines (V_k: the velocity
# magnitudes entering contact k), summed exactly by the oracle.  FP32
# rounding of kn, vn, delta^1.5 and the relative velocity is bounded relative
# to these terms.  Each test also prints the error against the net scale
#   sum_k |Fn_k + Ft_k| + m |g|,
# which is larger only where a contact's terms cancel (a separating contact
# whose damping offsets its Hertz repulsion, surfaces moving together).
FP32C_TOL = 1e-5
G = 9.80665


def world_step_magnitudes(orc, w, nsteps=1):
    """Step an oracle World; return the contact-force magnitude sums of its
    last step (net sums, with the term sums as .terms)."""
    if nsteps > 1:
        w.step(nsteps - 1)
    with orc.force_magnitudes(w.pos.shape[0]) as fm:
        w.step(1)
    return fm.copy()


def rel_error(got, ref, scale):
    """max over particles of |got - ref|_inf / scale (0 for no particles)."""
    got = np.asarray(got).reshape(len(scale), -1)
    ref = np.asarray(ref).reshape(len(scale), -1)
    if len(scale) == 0:
        return 0.0
    return float((np.abs(got - ref).max(1) / scale).max())


def force_errors(got, ref, fm, mass, radius=None):
    """Errors of per-particle forces (or torques, with radius) against the
    term scale ("term", the bar) and the net scale ("net", its maximum and
    the share of particles within FP32C_TOL of it)."""
    got = np.asarray(got).reshape(len(mass), -1)
    ref = np.asarray(ref).reshape(len(mass), -1)
    lev = 1.0 if radius is None else np.asarray(radius)
    d = np.abs(got - ref).max(1)
    st = (fm.terms + mass * G) * lev
    sn = (np.asarray(fm) + mass * G) * lev
    en = d / sn
    return {"term": float((d / st).max()) if len(d) else 0.0,
            "net": float(en.max()) if len(d) else 0.0,
            "net_ok_share": float((en <= FP32C_TOL).mean()) if len(d) else 1.0}


def acc_error(acc, w, fm):
    """Per-particle force m a of a World's step against the FP32c bar."""
    return force_errors(acc * w.mass[:, None], w.acc * w.mass[:, None], fm, w.mass)
