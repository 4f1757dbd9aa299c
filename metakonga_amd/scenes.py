"""Seeded scene generators for the DEM workloads (pure numpy, portable).

Random numbers come from a hand-written SplitMix64 (Steele, Lea & Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014), evaluated in
counter mode so results do not depend on numpy's generator versions.

Geometry follows the reference's definitions:
  plane::define(xw, pa, pc, pb) — source/xcar/plane.cpp:47-76 (4-point form:
      w2 = pa, w3 = pc, w4 = pb; l1 = |pa - xw|, u1 = (pa - xw)/l1, ...)
  plane::define(xw, pa, pb)     — source/xcar/plane.cpp:78-93 (3-point form,
      w2 = w3 = w4 = 0)
  cube::define(min, max)        — source/xcar/cube.cpp:55-75 (six 3-point planes)
  mass/inertia                  — source/simulation/dem_simulation.cpp:240-251
  contact property upload       — source/collision/contact.cpp:107-116 (rfric = 0)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 started at seed' = mix(seed, stream)."""
    base = _mix64(np.array([(seed * 0x100000001B3 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF],
                           dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        ctr = base + GOLDEN * (np.arange(1, n + 1, dtype=np.uint64))
    return _mix64(ctr)


def uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """U[0,1) doubles with 53 random bits."""
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


# --------------------------------------------------------------- materials
@dataclass
class MaterialSpec:
    density: float = 2500.0
    youngs: float = 1e7
    poisson: float = 0.3

    @property
    def shear(self) -> float:
        return self.youngs / (2.0 * (1.0 + self.poisson))


@dataclass
class ContactSpec:
    restitution: float = 0.8
    friction: float = 0.3
    rolling_friction: float = 0.0   # the reference uploads 0.0 (contact.cpp:112)
    cohesion: float = 0.0
    stiffness_ratio: float = 0.8    # DHS ks/kn; not given by the workload, stated


def contact_property(mi: MaterialSpec, mj: MaterialSpec, c: ContactSpec) -> abi.ContactProperty:
    """device_contact_property as uploaded by contact::cudaMemoryAlloc (contact.cpp:107-116)."""
    return abi.ContactProperty(mi.youngs, mj.youngs, mi.poisson, mj.poisson, mi.shear, mj.shear,
                               c.restitution, c.friction, c.rolling_friction, c.cohesion,
                               c.stiffness_ratio)


def mass_inertia(radius: np.ndarray, density: float):
    """dem_simulation.cpp:244-245: m = rho*(4/3)*pi*r^3, I = (2/5)*m*r^2."""
    m = density * (4.0 / 3.0) * math.pi * np.power(radius, 3.0)
    i = (2.0 / 5.0) * m * np.power(radius, 2.0)
    return m, i


# ---------------------------------------------------------------- geometry
def _v3(a):
    return np.asarray(a, dtype=np.float64)


def plane4(xw, pa, pc, pb) -> abi.PlaneInfo:
    """plane::define(xw, pa, pc, pb) — source/xcar/plane.cpp:47-76."""
    xw, pa, pc, pb = map(_v3, (xw, pa, pc, pb))
    w2, w3, w4 = pa.copy(), pc.copy(), pb.copy()
    a = pa - xw
    b = pb - xw
    l1 = math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    l2 = math.sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2])
    u1 = a / l1
    u2 = b / l2
    uw = np.array([u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                   u1[0] * u2[1] - u1[1] * u2[0]])
    return abi.PlaneInfo(l1, l2, abi._d3(u1), abi._d3(u2), abi._d3(uw), abi._d3(xw), abi._d3(a),
                         abi._d3(b), abi._d3(w2), abi._d3(w3), abi._d3(w4))


def plane3(xw, pa, pb) -> abi.PlaneInfo:
    """plane::define(xw, pa, pb) — source/xcar/plane.cpp:78-93 (w2 = w3 = w4 = 0)."""
    xw, pa, pb = map(_v3, (xw, pa, pb))
    a = pa - xw
    b = pb - xw
    l1 = math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    l2 = math.sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2])
    u1 = a / l1
    u2 = b / l2
    uw = np.array([u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                   u1[0] * u2[1] - u1[1] * u2[0]])
    z = abi.double3(0, 0, 0)
    return abi.PlaneInfo(l1, l2, abi._d3(u1), abi._d3(u2), abi._d3(uw), abi._d3(xw), abi._d3(a),
                         abi._d3(b), z, z, z)


def cube_planes(mn, mx) -> list:
    """cube::define(min, max) — source/xcar/cube.cpp:55-75."""
    mn, mx = _v3(mn), _v3(mx)
    sx = float(np.linalg.norm(mx - np.array([mn[0], mx[1], mx[2]])))
    sy = float(np.linalg.norm(mx - np.array([mx[0], mn[1], mx[2]])))
    sz = float(np.linalg.norm(mx - np.array([mx[0], mx[1], mn[2]])))
    V = lambda x, y, z: np.array([x, y, z])
    return [
        plane3(mn, mn + V(0, 0, sz), mn + V(sx, 0, 0)),
        plane3(mn, mn + V(0, sy, 0), mn + V(0, 0, sz)),
        plane3(mn + V(sx, 0, 0), mn + V(sx, 0, sz), mn + V(sx, sy, 0)),
        plane3(mn, mn + V(sx, 0, 0), mn + V(0, sy, 0)),
        plane3(mn + V(0, 0, sz), mn + V(0, sy, sz), mn + V(sx, 0, sz)),
        plane3(mn + V(0, sy, 0), mn + V(sx, sy, 0), mn + V(0, sy, sz)),
    ]


def open_box_planes(mn, mx) -> list:
    """Floor and four side walls (open top, y up), 4-point planes so edges and
    corners are valid."""
    x0, y0, z0 = mn
    x1, y1, z1 = mx
    return [
        plane4((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # floor
        plane4((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)),  # x = x0
        plane4((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)),  # x = x1
        plane4((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)),  # z = z0
        plane4((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # z = z1
    ]


def circumcenter(P, Q, R):
    """Circumcentre as numeric::utility::calculate_center_of_triangle
    (source/numeric_utility.cpp:20-45), vectorised over rows."""
    P, Q, R = (np.asarray(x, dtype=np.float64) for x in (P, Q, R))
    V = Q - P
    W = R - P
    N = np.cross(V, W)
    N = N / np.sqrt((N * N).sum(-1, keepdims=True))
    M1 = 0.5 * (Q + P)
    M2 = 0.5 * (R + P)
    D1 = np.cross(N, V)
    D2 = np.cross(N, W)
    den_xy = D1[..., 0] * D2[..., 1] - D1[..., 1] * D2[..., 0]
    den_xz = D1[..., 0] * D2[..., 2] - D1[..., 2] * D2[..., 0]
    den_yz = D1[..., 1] * D2[..., 2] - D1[..., 2] * D2[..., 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        t_xy = (D2[..., 0] * (M1[..., 1] - M2[..., 1])) / den_xy - (D2[..., 1] * (M1[..., 0] - M2[..., 0])) / den_xy
        t_xz = (D2[..., 0] * (M1[..., 2] - M2[..., 2])) / den_xz - (D2[..., 2] * (M1[..., 0] - M2[..., 0])) / den_xz
        t_yz = (D2[..., 1] * (M1[..., 2] - M2[..., 2])) / den_yz - (D2[..., 2] * (M1[..., 1] - M2[..., 1])) / den_yz
    t = np.where(np.abs(den_xy) > 1e-13, t_xy,
                 np.where(np.abs(den_xz) > 1e-13, t_xz, np.where(np.abs(den_yz) > 1e-13, t_yz, 0.0)))
    return M1 + t[..., None] * D1


def motion(velocity=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0), start_time=0.0,
           end_time=0.0) -> abi.Motion:
    """A prescribed rigid motion (xdem_motion): constant velocity and angular
    velocity from start_time (until end_time when end_time > start_time)."""
    return abi.Motion(abi._d3(velocity), abi._d3(omega), float(start_time), float(end_time))


def drum_info(origin=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0), radius=0.5, half_length=0.2,
              omega=0.0, caps=True) -> abi.DrumInfo:
    """A rotating drum wall (xdem_drum_info); the axis is normalised here."""
    a = _v3(axis)
    a = a / math.sqrt(float(a @ a))
    return abi.DrumInfo(abi._d3(origin), abi._d3(a), float(radius), float(half_length),
                        float(omega), 1 if caps else 0, 0)


@dataclass
class PolygonObject:
    """One rigid triangle-mesh object (device_polygon_mass_info): pose of
    its frame and an optional prescribed motion about its origin."""
    origin: tuple = (0.0, 0.0, 0.0)
    ep: tuple = (1.0, 0.0, 0.0, 0.0)       # Euler parameters, e0 first
    motion: abi.Motion | None = None


# ------------------------------------------------------------------ scenes
@dataclass
class Scene:
    """One DEM scene in the reference data model (host arrays)."""
    name: str
    pos: np.ndarray                 # (N,4) x,y,z,r float64
    dt: float
    nsteps: int
    gravity: tuple = (0.0, -9.80665, 0.0)   # source/xcar/model.cpp:5
    world_origin: tuple = (-1.0, -1.0, -1.0)
    cell_size: float = 0.0
    grid_size: tuple = (128, 128, 128)
    material: MaterialSpec = field(default_factory=MaterialSpec)
    contact: ContactSpec = field(default_factory=ContactSpec)
    planes: list = field(default_factory=list)        # PlaneInfo, applied in order
    cylinders: list = field(default_factory=list)     # CylinderInfo
    tri: np.ndarray | None = None                     # (T,3,3) triangle vertices
    tri_object: np.ndarray | None = None              # (T,) object of each triangle (default 0)
    objects: list = field(default_factory=list)       # PolygonObject per object id
    drums: list = field(default_factory=list)         # DrumInfo, applied after the cylinders
    plane_motions: dict = field(default_factory=dict)  # plane index -> abi.Motion
    rolling: bool = True
    tcm: int = 1
    seed: int = 0
    tangential_history: bool = False   # fused path: DHS with a tangential spring per contact
    # real-time release (FP64 simulation): (steps between groups, particles per group or
    # a list of group sizes); particles join in index order, the first group at step 1
    # (dem_simulation.cpp:449-450); release_period gives the steps between groups
    release: tuple | None = None

    @property
    def np(self) -> int:
        return int(self.pos.shape[0])

    def default_cell_size(self) -> float:
        """2 max(radius, triangle circumradius) (dem_simulation.cpp:226-227,249-250,261)."""
        rmax = float(self.pos[:, 3].max())
        if self.tri is not None:
            rmax = max(rmax, float(self.polygon_arrays()[1][:, 3].max()))
        return 2.0 * rmax

    def params(self) -> abi.DeviceParameters:
        cs = self.cell_size or self.default_cell_size()
        gs = self.grid_size
        d = abi.DeviceParameters()
        d.rollingCondition = self.rolling
        d.np = self.np
        d.nsphere = 0
        d.ncell = gs[0] * gs[1] * gs[2]
        d.grid_size = abi.uint3(*gs)
        d.dt = self.dt
        d.half2dt = 0.5 * self.dt * self.dt
        d.cell_size = cs
        d.cohesion = 0.0
        d.gravity = abi._d3(self.gravity)
        d.world_origin = abi._d3(self.world_origin)
        return d

    def polygon_arrays(self):
        """(triangles, circumspheres (T,4), objects) as the C ABI takes them
        (contact_particles_polygonObjects::define, source/contact_particles_polygonObjects.cpp:47-123:
        circumcentre and circumradius per triangle)."""
        tri = np.ascontiguousarray(self.tri, dtype=np.float64)
        nt = tri.shape[0]
        ids = (np.zeros(nt, np.int32) if self.tri_object is None
               else np.asarray(self.tri_object, np.int32))
        objs = self.objects or [PolygonObject()]
        cc = circumcenter(tri[:, 0], tri[:, 1], tri[:, 2])
        rad = np.sqrt(((cc - tri[:, 0]) ** 2).sum(1))
        sph = np.ascontiguousarray(np.concatenate([cc, rad[:, None]], 1))
        dpi = abi.struct_array(abi.PolygonInfo, [
            abi.PolygonInfo(int(ids[t]), abi._d3(tri[t, 0]), abi._d3(tri[t, 1]), abi._d3(tri[t, 2]))
            for t in range(nt)])
        dpmi = []
        for ob in objs:
            mi = abi.PolygonMassInfo()
            mi.origin = abi._d3(ob.origin)
            mi.ep = abi.double4(*ob.ep)
            dpmi.append(mi)
        return dpi, sph, abi.struct_array(abi.PolygonMassInfo, dpmi)

    def contact_pp(self) -> abi.ContactProperty:
        return contact_property(self.material, self.material, self.contact)

    def mass_inertia(self):
        return mass_inertia(self.pos[:, 3], self.material.density)


def _pow2_at_least(n: int) -> int:
    p = 1
    while p < n:
        p <<= 1
    return p


def release_period(dt: float, time_per_group: float | None = None,
                   per_second: int | None = None) -> int:
    """Steps between real-time releases, dem_simulation::initialize
    (dem_simulation.cpp:210-219): one group every time_per_group seconds gives
    (unsigned)((time_per_group + 1e-9) / dt); one particle at a time at
    per_second particles a second gives (unsigned)((1.0 / per_second) / dt).
    0 means no release (every particle is simulated from the start)."""
    if (time_per_group is None) == (per_second is None):
        raise ValueError("give time_per_group or per_second")
    if per_second is not None:
        if int(per_second) <= 0:
            raise ValueError("per_second must be a positive count")
        return int((1.0 / int(per_second)) / float(dt))
    return int((float(time_per_group) + 1e-9) / float(dt))


def as_f32_exact(x: np.ndarray) -> np.ndarray:
    """Round to the nearest FP32 value and widen: such values are identical in
    the FP64 and the FP32 state representations."""
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def settling_box(n_side=20, radius=5e-3, box=0.3, seed=0xD3E0, dt=1e-5, nsteps=2000,
                 compat_grid=False, name="C0-settling") -> Scene:
    """C0: n_side^3 monodisperse spheres on a jittered cubic lattice (pitch 2.4 r,
    jitter U(-0.2r, 0.2r) per axis, zero velocity) above the floor of an open
    box of 5 planes; cell size 2r.

    Grid: fitted (wo = box_min - cs, 32^3 for the 0.3 m box), or the reference's
    compat grid (wo = -1, 128^3, box placed inside [-1, 0.28]^3)."""
    r = float(as_f32_exact(radius))
    pitch = 2.4 * r
    bmin = np.array([-box / 2, -0.6, -box / 2])
    bmax = bmin + box
    idx = np.arange(n_side ** 3)
    ix, iy, iz = idx % n_side, (idx // n_side) % n_side, idx // (n_side * n_side)
    span = (n_side - 1) * pitch
    cx = bmin[0] + box / 2 - span / 2 + ix * pitch
    cz = bmin[2] + box / 2 - span / 2 + iz * pitch
    cy = bmin[1] + 2.4 * r + iy * pitch
    j = (uniform(seed, 3 * idx.size, 1).reshape(-1, 3) * 2.0 - 1.0) * 0.2 * r
    pos = np.stack([cx + j[:, 0], cy + j[:, 1], cz + j[:, 2], np.full(idx.size, r)], axis=1)
    cs = 2.0 * r
    if compat_grid:
        wo, gs = (-1.0, -1.0, -1.0), (128, 128, 128)
    else:
        wo = tuple(bmin - cs)
        g = _pow2_at_least(int(math.ceil(round((box + 2 * cs) / cs, 6))))
        gs = (g, g, g)
    return Scene(name=name, pos=np.ascontiguousarray(pos), dt=dt, nsteps=nsteps,
                 world_origin=wo, cell_size=cs, grid_size=gs,
                 planes=open_box_planes(bmin, bmax), seed=seed)


def hopper(n=1_000_000, r_min=2e-3, r_max=4e-3, seed=0xD3E1, dt=5e-6, nsteps=5000,
           slot=0.03, name="C1-hopper") -> Scene:
    """C1: polydisperse spheres (radii U[r_min, r_max], rounded to FP32 values)
    in a flat-bottomed hopper of 6 planes: 4 side walls and two bottom plates
    inclined towards a slot of width `slot` along z.  Seeded random
    non-overlapping fill: a cubic lattice of pitch 2.1 r_max, each sphere
    jittered inside its own lattice cell (no overlap by construction).  Cell
    size 2 r_max, grid 128^3 (1.024 m at 8 mm cells)."""
    pitch = 2.1 * r_max
    width = max(0.12, 0.9 * math.sqrt(n / 1.0e6))
    per_row = int((width - 2 * (r_max + 1e-3)) / pitch)
    per_layer = per_row * per_row
    layers = int(math.ceil(n / per_layer))
    y_b = 0.1 * width / 0.9
    y0 = y_b + 1.5 * r_max
    idx = np.arange(n)
    ix, iz, iy = idx % per_row, (idx // per_row) % per_row, idx // per_layer
    x0 = -(per_row - 1) * pitch / 2
    rad = as_f32_exact(r_min + (r_max - r_min) * uniform(seed, n, 1))
    slack = (pitch / 2 - rad) * 0.95
    j = uniform(seed, 3 * n, 2).reshape(-1, 3) * 2.0 - 1.0
    pos = np.stack([x0 + ix * pitch + j[:, 0] * slack,
                    y0 + iy * pitch + j[:, 1] * slack,
                    x0 + iz * pitch + j[:, 2] * slack, rad], axis=1)
    H = y0 + layers * pitch + 0.05
    hw = width / 2
    s = slot / 2
    planes = [
        plane4((-hw, y_b, -hw), (-hw, H, -hw), (-hw, H, hw), (-hw, y_b, hw)),   # x = -hw
        plane4((hw, y_b, -hw), (hw, H, -hw), (hw, H, hw), (hw, y_b, hw)),       # x = +hw
        plane4((-hw, 0.0, -hw), (hw, 0.0, -hw), (hw, H, -hw), (-hw, H, -hw)),   # z = -hw
        plane4((-hw, 0.0, hw), (hw, 0.0, hw), (hw, H, hw), (-hw, H, hw)),       # z = +hw
        plane4((-hw, y_b, -hw), (-s, 0.0, -hw), (-s, 0.0, hw), (-hw, y_b, hw)),  # left plate
        plane4((s, 0.0, -hw), (hw, y_b, -hw), (hw, y_b, hw), (s, 0.0, hw)),      # right plate
    ]
    cs = 2.0 * r_max
    g = max(_pow2_at_least(int(math.ceil((width + 4 * cs) / cs))),
            _pow2_at_least(int(math.ceil((H + 4 * cs) / cs))))
    wo = (-g * cs / 2, -2 * cs - 0.0, -g * cs / 2)
    return Scene(name=name, pos=np.ascontiguousarray(pos), dt=dt, nsteps=nsteps,
                 world_origin=wo, cell_size=cs, grid_size=(g, g, g), planes=planes, seed=seed)


def dense_bed(n=1_000_000, r=4e-3, seed=0xD3E5, dt=5e-6, nsteps=1000, squeeze=0.985, poly=0.02,
              name="dense-bed") -> Scene:
    """Contact-rich compressed packing (a benchmark state, not a reference
    config): a face-centred cubic lattice with nearest-neighbour distance
    d = squeeze * 2r (4 sites per cubic cell of side d * sqrt(2)), filled in
    site order (x, then z, then y layers) in a closed 6-plane box whose walls
    sit squeeze * r outside the outermost sites; radii U[(1 - poly) r, r]
    rounded to FP32 values.  Neighbours whose radii sum exceeds d overlap by at
    most (1 - squeeze) * 2r: with the defaults about 7/8 of the 12 lattice
    neighbours touch, solid fraction about 0.7.  Cell size 2r."""
    r = float(as_f32_exact(r))
    d = squeeze * 2.0 * r
    a = d * math.sqrt(2.0)
    m = int(math.ceil((n / 4.0) ** (1.0 / 3.0)))
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    c = np.arange(m)
    # cells in (y, z, x) order, x fastest; 4 sites per cell
    cy, cz, cx = np.meshgrid(c, c, c, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1).astype(np.float64)
    sites = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)[:n] * a
    rad = as_f32_exact(r * (1.0 - poly * uniform(seed, n, 1)))
    lo = -squeeze * r
    hi = (m - 0.5) * a + squeeze * r
    top = float(sites[:, 1].max()) + squeeze * r
    off = 2.0 * r  # box corner at (off, off, off) - lo
    pos = np.concatenate([sites + (off - lo), rad[:, None]], 1)
    mn = (off, off, off)
    mx = (off + hi - lo, off + top - lo, off + hi - lo)
    cs = 2.0 * r
    g = _pow2_at_least(int(math.ceil((max(mx) + 2.0 * r) / cs)) + 2)
    planes = cube_planes(mn, mx)
    return Scene(name=name, pos=np.ascontiguousarray(pos), dt=dt, nsteps=nsteps,
                 world_origin=(0.0, 0.0, 0.0), cell_size=cs, grid_size=(g, g, g), planes=planes,
                 seed=seed)


def drum(n=4_000_000, radius=3e-3, R=0.5, L=1.6, rpm=30.0, ntri=20000, seed=0xD3E2, dt=5e-6,
         nsteps=2000, name="C2-drum", grid=(256, 256, 256)) -> Scene:
    """C2: n spheres of radius r in a drum (inner shell and end caps, axis z,
    spinning at rpm) with a seeded random triangle-mesh baffle of ntri
    triangles that turns with the drum: a radial lifter plate in the plane
    y = 0 from 0.8 R to the shell along the whole length, a grid of square
    cells of side h (ntri = 2 x cells; h about 4 mm for the defaults) split
    along a random diagonal each, interior vertices jittered by 0.025 h in the
    plate (circumradius <= r, which keeps the cell size at 2 r).  L = 1.6 m
    instead of 0.4 m: 4e6 spheres of r = 3 mm (0.452 m^3 solid) cannot fit a
    0.314 m^3 drum.  The fill is the lowest part of a jittered lattice (pitch
    2.1 r) in every cross-section, clear of the plate by 2.5 r; gravity -y.
    The default grid (256^3 cells of 2 max(r, circumradius)) wraps once along
    the axis: hashed cells alias, which the distance test keeps correct."""
    r = float(as_f32_exact(radius))
    pitch = 2.1 * r
    omega = rpm * 2.0 * math.pi / 60.0
    # baffle: nr x nz square cells over [0.8 R, R] x [-L/2, L/2] in the plane y = 0
    w = 0.2 * R
    nr = max(1, int(round(math.sqrt(ntri / 2 * w / L))))
    nzc = max(1, ntri // (2 * nr))
    hr, hz = w / nr, L / nzc
    vr = 0.8 * R + hr * np.arange(nr + 1)
    vz = -L / 2 + hz * np.arange(nzc + 1)
    VR, VZ = np.meshgrid(vr, vz, indexing="ij")
    jit = (uniform(seed, 2 * VR.size, 3).reshape(VR.shape + (2,)) * 2 - 1) * 0.025
    VR[1:-1, :] += jit[1:-1, :, 0] * hr
    VZ[:, 1:-1] += jit[:, 1:-1, 1] * hz
    V = np.stack([VR, np.zeros_like(VR), VZ], -1)
    a, b_, c_, d_ = V[:-1, :-1], V[1:, :-1], V[1:, 1:], V[:-1, 1:]
    flip = (uniform(seed, nr * nzc, 4).reshape(nr, nzc) < 0.5)[..., None]
    t1 = np.stack([a, b_, np.where(flip, d_, c_)], -2)
    t2 = np.stack([np.where(flip, b_, a), c_, d_], -2)
    tri = np.concatenate([t1.reshape(-1, 3, 3), t2.reshape(-1, 3, 3)], 0)
    # lattice over the cross-section, inside the shell, clear of the plate
    per = int((2 * R - 4 * r) / pitch)
    xs = (np.arange(per) - (per - 1) / 2) * pitch
    yy, xx = np.meshgrid(xs, xs, indexing="ij")
    inside = ((xx ** 2 + yy ** 2) < (R - 2 * r) ** 2) & ~(
        (xx > 0.8 * R - 2.5 * r) & (np.abs(yy) < 2.5 * r))
    yx = np.stack([xx[inside], yy[inside]], 1)
    yx = yx[np.lexsort((yx[:, 0], yx[:, 1]))]          # lowest first
    nz = int((L - 4 * r) / pitch)
    m = min(int(math.ceil(n / nz)), yx.shape[0])
    idx = np.arange(min(n, m * nz))
    sl, k = idx // m, idx % m
    j = (uniform(seed, 3 * idx.size, 1).reshape(-1, 3) * 2 - 1) * (pitch / 2 - r) * 0.9
    pos = np.stack([yx[k, 0] + j[:, 0], yx[k, 1] + j[:, 1], -L / 2 + 2 * r + sl * pitch + j[:, 2],
                    np.full(idx.size, r)], 1)
    sc = Scene(name=name, pos=np.ascontiguousarray(pos), dt=dt, nsteps=nsteps, tri=tri, seed=seed,
               objects=[PolygonObject(origin=(0.0, 0.0, 0.0),
                                      motion=motion(omega=(0.0, 0.0, omega)))],
               drums=[drum_info((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), R, L / 2, omega, True)],
               grid_size=grid)
    cs = sc.default_cell_size()
    sc.cell_size = cs
    sc.world_origin = (-grid[0] * cs / 2, -grid[1] * cs / 2, -grid[2] * cs / 2)
    return sc


def log_uniform_counts(seed: int, k: int, lo: int, hi: int, stream: int = 7) -> np.ndarray:
    u = uniform(seed, k, stream)
    return np.floor(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u)).astype(np.int64)


def batch_scenes(nscene=256, lo=2000, hi=200_000, seed=0xD3E3, dt=5e-6, nsteps=1000):
    """C3: independent box-settling scenes, N_s ~ logU[lo, hi], one radius per
    scene r_s ~ U[2, 5] mm (monodisperse within a scene)."""
    counts = log_uniform_counts(seed, nscene, lo, hi)
    radii = as_f32_exact(2e-3 + 3e-3 * uniform(seed, nscene, 8))
    out = []
    for s in range(nscene):
        out.append(box_scene(int(counts[s]), float(radii[s]), seed + 1000 + s, dt, nsteps,
                             name=f"C3-scene{s}"))
    return out


def box_scene(n: int, r: float, seed: int, dt: float, nsteps: int, name="box") -> Scene:
    """An n-sphere jittered lattice (pitch 2.4 r, jitter 0.2 r) above the floor of
    an open box sized to the lattice; fitted power-of-two grid, cell 2r."""
    r = float(as_f32_exact(r))
    pitch = 2.4 * r
    nx = max(2, int(math.ceil(n ** (1.0 / 3.0))))
    ny = int(math.ceil(n / (nx * nx)))
    width = nx * pitch + 2 * r
    height = ny * pitch * 1.5 + 4 * r
    bmin = np.array([0.0, 0.0, 0.0])
    bmax = np.array([width, height, width])
    idx = np.arange(n)
    ix, iz, iy = idx % nx, (idx // nx) % nx, idx // (nx * nx)
    j = (uniform(seed, 3 * n, 1).reshape(-1, 3) * 2 - 1) * 0.2 * r
    pos = np.stack([r + pitch / 2 + ix * pitch + j[:, 0] - r / 2,
                    2.4 * r + iy * pitch + j[:, 1],
                    r + pitch / 2 + iz * pitch + j[:, 2] - r / 2, np.full(n, r)], 1)
    cs = 2.0 * r
    gxz = _pow2_at_least(int(math.ceil(width / cs)) + 3)
    gy = _pow2_at_least(int(math.ceil(height / cs)) + 3)
    wo = (-cs, -cs, -cs)
    return Scene(name=name, pos=np.ascontiguousarray(pos), dt=dt, nsteps=nsteps, world_origin=wo,
                 cell_size=cs, grid_size=(gxz, gy, gxz), planes=open_box_planes(bmin, bmax),
                 seed=seed)
