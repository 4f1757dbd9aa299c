"""Python mirror of the reference's host objects over the C ABI (libxdem.so).

`Simulation` follows dem_simulation (source/simulation/dem_simulation.cpp:204-325,
447-469): walls are applied in the order they are added, like the reference's
contact map; `Grid` follows grid_base/neighborhood_cell (include/grid_base.h).
These are thin ctypes wrappers: every computation runs in libxdem.so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi


def _desc(scene) -> abi.SimDesc:
    d = abi.SimDesc()
    d.np = scene.np
    d.dt = scene.dt
    d.gravity = abi._d3(scene.gravity)
    d.world_origin = abi._d3(scene.world_origin)
    d.cell_size = scene.cell_size
    d.grid_size = abi.uint3(*scene.grid_size)
    d.rolling = scene.rolling
    d.tcm = scene.tcm
    m = scene.material
    d.particle = abi.Material(m.density, m.youngs, m.poisson, m.shear)
    c = scene.contact
    d.pp = abi.ContactPairDesc(c.restitution, c.friction, c.rolling_friction, c.cohesion,
                               c.stiffness_ratio, d.particle)
    return d


class Simulation:
    """One scene on the plain FP64 GPU path (xdem_sim_*)."""

    def __init__(self, scene, lib=None):
        self.lib = lib or abi.load()
        self.scene = scene
        self.desc = _desc(scene)
        h = C.c_void_p()
        pos = abi.as_f64(scene.pos)
        abi.check(self.lib, self.lib.xdem_sim_create(C.byref(self.desc), abi.ptr(pos), C.byref(h)))
        self.h = h
        c = scene.contact
        wall = abi.ContactPairDesc(c.restitution, c.friction, c.rolling_friction, c.cohesion,
                                   c.stiffness_ratio, self.desc.particle)
        for pl in scene.planes:
            abi.check(self.lib, self.lib.xdem_sim_add_plane(h, C.byref(pl), C.byref(wall)))
        for cy in scene.cylinders:
            abi.check(self.lib, self.lib.xdem_sim_add_cylinder(h, C.byref(cy), C.byref(wall)))
        for dr in scene.drums:
            abi.check(self.lib, self.lib.xdem_sim_add_drum(h, C.byref(dr), C.byref(wall)))
        self.ntri = self.nobj = 0
        if scene.tri is not None:
            dpi, sph, dpmi = scene.polygon_arrays()
            per = (abi.ContactPairDesc * len(dpmi))(*([wall] * len(dpmi)))
            abi.check(self.lib, self.lib.xdem_sim_set_polygons(
                h, C.cast(dpi, C.c_void_p), abi.ptr(sph), len(dpi), C.cast(dpmi, C.c_void_p),
                C.cast(per, C.c_void_p), len(dpmi)))
            self.ntri, self.nobj = len(dpi), len(dpmi)
        # contact_particles_plane::cuda_collision re-uploads a moving plane every
        # step (source/collision/contact_particles_plane.cpp:33-49)
        for i, m in sorted(scene.plane_motions.items()):
            abi.check(self.lib, self.lib.xdem_sim_set_motion(h, abi.OBJECT_PLANE, i, C.byref(m)))
        for o, ob in enumerate(scene.objects):
            if ob.motion is not None:
                abi.check(self.lib, self.lib.xdem_sim_set_motion(h, abi.OBJECT_POLYGON, o,
                                                                  C.byref(ob.motion)))
        if scene.release is not None and int(scene.release[0]) > 0:
            every, group = scene.release
            if np.ndim(group) == 0:
                abi.check(self.lib, self.lib.xdem_sim_set_release(h, int(every), int(group)))
            else:
                g = np.ascontiguousarray(group, dtype=np.uint32)
                abi.check(self.lib, self.lib.xdem_sim_set_release_groups(
                    h, int(every), abi.ptr(g), len(g)))

    @property
    def np(self):
        return int(self.lib.xdem_sim_np(self.h))

    @property
    def active(self):
        """Particles simulated so far (all of them unless a release is set)."""
        return int(self.lib.xdem_sim_active(self.h))

    def step(self, n=1):
        abi.check(self.lib, self.lib.xdem_sim_step(self.h, n))

    def sync(self):
        abi.check(self.lib, self.lib.xdem_synchronize())

    def state(self):
        n = self.np
        out = dict(pos=np.empty((n, 4)), vel=np.empty((n, 3)), omega=np.empty((n, 3)),
                   acc=np.empty((n, 3)), alpha=np.empty((n, 3)), force=np.empty((n, 3)),
                   moment=np.empty((n, 3)))
        abi.check(self.lib, self.lib.xdem_sim_get_state(
            self.h, *(abi.ptr(out[k]) for k in ("pos", "vel", "omega", "acc", "alpha", "force",
                                                 "moment"))))
        return out

    def set_state(self, **kw):
        n = self.np
        arrs = []
        for k, shp in (("pos", 4), ("vel", 3), ("omega", 3), ("acc", 3), ("alpha", 3),
                       ("force", 3), ("moment", 3)):
            a = kw.get(k)
            arrs.append(None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(n, shp))
        abi.check(self.lib, self.lib.xdem_sim_set_state(self.h, *(abi.ptr(a) for a in arrs)))

    @property
    def steps(self):
        return int(self.lib.xdem_sim_steps(self.h))

    def polygons(self):
        """Current triangles (ctypes array of PolygonInfo) and circumspheres (T,4)."""
        tris = (abi.PolygonInfo * self.ntri)()
        sph = np.empty((self.ntri, 4))
        abi.check(self.lib, self.lib.xdem_sim_get_polygons(self.h, C.cast(tris, C.c_void_p),
                                                            abi.ptr(sph)))
        return tris, sph

    def set_object_state(self, obj, origin, ep, vel, omega):
        """Pose and velocity of polygon object `obj` from a multibody solver;
        its triangles are re-posed before the next step."""
        st = abi.PolygonMassInfo()
        st.origin = abi._d3(origin)
        st.ep = abi.double4(*ep)
        st.vel = abi._d3(vel)
        st.omega = abi._d3(omega)
        abi.check(self.lib, self.lib.xdem_sim_set_object_state(self.h, obj, C.byref(st)))

    def objects(self):
        """Polygon objects (PolygonMassInfo): pose, velocity, last reaction."""
        objs = (abi.PolygonMassInfo * self.nobj)()
        abi.check(self.lib, self.lib.xdem_sim_get_objects(self.h, C.cast(objs, C.c_void_p)))
        return objs

    # state files (include/xdem_io.h)
    def save_part(self, directory, part):
        abi.check(self.lib, self.lib.xdem_sim_save_part(self.h, os.fsencode(directory), part))

    def save_final(self, path):
        abi.check(self.lib, self.lib.xdem_sim_save_final(self.h, os.fsencode(path)))

    def load_bfr(self, path):
        abi.check(self.lib, self.lib.xdem_sim_load_bfr(self.h, os.fsencode(path)))

    def save_checkpoint(self, path):
        abi.check(self.lib, self.lib.xdem_sim_save_checkpoint(self.h, os.fsencode(path)))

    def load_checkpoint(self, path):
        abi.check(self.lib, self.lib.xdem_sim_load_checkpoint(self.h, os.fsencode(path)))

    def mass_inertia(self):
        ptrs = [C.c_void_p() for _ in range(9)]
        abi.check(self.lib, self.lib.xdem_sim_device_arrays(self.h, *(C.byref(p) for p in ptrs)))
        n = self.np
        m = np.empty(n)
        i = np.empty(n)
        abi.check(self.lib, self.lib.xdem_memcpy_dtoh(abi.ptr(m), ptrs[7], 8 * n))
        abi.check(self.lib, self.lib.xdem_memcpy_dtoh(abi.ptr(i), ptrs[8], 8 * n))
        return m, i

    def close(self):
        if self.h:
            self.lib.xdem_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
