"""Model input: the reference's particle generators, triangle-mesh importers and
the .xdm text model, producing a `scenes.Scene` for the GPU paths.

* Generators restate particleManager::CreateCubeParticle / CreatePlaneParticle
  / CreateCircleParticle and the calculateNum*Particles counts
  (source/particleManager.cpp:170-618).  The reference jitters with
  frand() = rand() / (float)RAND_MAX seeded by GetTickCount (:368, :560); here
  the draws come from the seeded splitmix64 stream of `scenes.uniform`, rounded
  to FP32 like frand().  The layouts, spacings and counts are the reference's.
* Importers restate vpolygon::_loadMS3DASCII / _loadSTLASCII
  (source/xcar/vpolygon.cpp:97-190,190-300): vertex coordinates in mm scaled by
  0.001, MilkShape vertex/normal/triangle tables, STL facets; the object frame
  follows polygonObject::define (source/xcar/polygonObject.cpp:45-160): a
  MilkShape object sits at the mean of its triangles' circumcentres, an STL
  object at its given location.
* The .xdm parser follows modelManager::OpenModel
  (source/xcar/modelManager.cpp:55-107) and the section readers it calls:
  dem_model::Open (source/dem_model.cpp:40-48), particleManager::Open
  (source/particleManager.cpp:42-130), geometryObjects::Open
  (source/xcar/geometryObjects.cpp:26-96) and contactManager::Open
  (source/contactManager.cpp:40-63).  MULTIBODY_MODEL_DATA (the MBD model) is
  recorded as present and otherwise skipped.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi, scenes


def _frand(seed: int, n: int, stream: int) -> np.ndarray:
    return scenes.uniform(seed, n, stream).astype(np.float32).astype(np.float64)


# ----------------------------------------------------------------- counts
def num_cube_particles(dx, dy, dz, min_radius, max_radius) -> int:
    """calculateNumCubeParticles (:200-211): (floor(dim / 2r) - 1) per axis
    (monodisperse only, as the reference)."""
    if min_radius != max_radius:
        raise ValueError("the cube generator is monodisperse (min_radius == max_radius)")
    d = 2.0 * min_radius
    return _ndim(dx, d) * _ndim(dy, d) * _ndim(dz, d)


def _ndim(length, d):
    """(length / d).To<unsigned int>() - 1, floored at 0."""
    return max(int(length / d) - 1, 0)


def num_plane_particles(dx, ny, dz, min_radius, max_radius) -> int:
    """calculateNumPlaneParticles (:178-188)."""
    d = 2.0 * max_radius
    return _ndim(dx, d) * int(ny) * _ndim(dz, d)


def _circle_rings(dia, r):
    cr = 0.5 * dia
    nr = int(cr / (2.0 * r)) - 1
    if nr < 0:
        return [], 0.0
    rr = 2.0 * r * nr + r
    space = (cr - rr) / (nr + 1)
    rings = []
    for i in range(1, nr + 1):
        _r = i * (2.0 * r + space)
        dth = (2.0 * r + space) / _r
        rings.append((_r, int((2.0 * math.pi) / dth)))
    return rings, space


def num_circle_particles(dia, ny, min_radius, max_radius) -> int:
    """calculateNumCircleParticles (:190-205): a centre particle plus
    floor(2 pi / dth) per ring, per layer."""
    rings, _ = _circle_rings(dia, max_radius)
    return (1 + sum(n for _, n in rings)) * int(ny)


# ------------------------------------------------------------- generators
def cube_particles(dim, start, radius, seed=0):
    """CreateCubeParticle (:213-296), monodisperse: an (nx, ny, nz) lattice of
    pitch 2r + space, space = (L - 2r n) / (n + 1) per axis, starting at
    start + r + space, jittered by r * space * frand() per axis; z outer, x inner.
    Returns (N, 4) x, y, z, r."""
    dx, dy, dz = (float(v) for v in dim)
    lx, ly, lz = (float(v) for v in start)
    r = float(radius)
    d = 2.0 * r
    n = [_ndim(dx, d), _ndim(dy, d), _ndim(dz, d)]
    dimv = np.array([dx, dy, dz])
    space = (dimv - d * np.array(n, float)) / (np.array(n, float) + 1)
    gab = d + space
    ran = r * space
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    N = x.size
    f = _frand(seed, 3 * N, 11).reshape(N, 3)
    out = np.empty((N, 4))
    out[:, 0] = lx + r + space[0] + x * gab[0] + ran[0] * f[:, 0]
    out[:, 1] = ly + r + space[1] + y * gab[1] + ran[1] * f[:, 1]
    out[:, 2] = lz + r + space[2] + z * gab[2] + ran[2] * f[:, 2]
    out[:, 3] = r
    return out


def _radii(pos, min_radius, max_radius, seed, stream):
    if min_radius != max_radius:
        pos[:, 3] = min_radius + (max_radius - min_radius) * _frand(seed, pos.shape[0], stream)
    return pos


def plane_particles(dx, ny, dz, start, min_radius, max_radius, seed=0):
    """CreatePlaneParticle (:298-418): ny groups of an (nx, nz) layer at height
    ly (the reference releases one group at a time), pitch 2r + space with
    r = max_radius, jitter 100 r space frand() in x and z; radii U[min, max]
    when they differ.  Returns ((N, 4) positions, group sizes)."""
    lx, ly, lz = (float(v) for v in start)
    r = float(max_radius)
    d = 2.0 * r
    nx, nz = _ndim(dx, d), _ndim(dz, d)
    sx = (dx - d * nx) / (nx + 1)
    sz = (dz - d * nz) / (nz + 1)
    gx, gz = d + sx, d + sz
    ranx, ranz = 100.0 * r * sx, 100.0 * r * sz
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    xx, zz = xx.ravel(), zz.ravel()
    per = xx.size
    N = per * int(ny)
    f = _frand(seed, 2 * N, 12).reshape(N, 2)
    pos = np.empty((N, 4))
    pos[:, 0] = lx + sx + np.tile(xx, ny) * gx + ranx * f[:, 0]
    pos[:, 1] = ly
    pos[:, 2] = lz + sz + np.tile(zz, ny) * gz + ranz * f[:, 1]
    pos[:, 3] = r
    return _radii(pos, min_radius, max_radius, seed, 13), [per] * int(ny)


def circle_particles(dia, ny, start, min_radius, max_radius, seed=0):
    """CreateCircleParticle (:420-618): per layer a centre particle and rings
    of radius i (2r + space), i = 1..nr, with floor(2 pi / dth) particles at
    angles j * 2 pi / npr, all at height ly; radii U[min, max] when they
    differ.  Returns ((N, 4) positions, ring particles per layer)."""
    lx, ly, lz = (float(v) for v in start)
    r = float(max_radius)
    rings, _ = _circle_rings(dia, r)
    layer = [(lx, ly, lz)]
    for _r, npr in rings:
        dth = (2.0 * math.pi) / npr
        for j in range(npr):
            layer.append((lx + _r * math.cos(dth * j), ly, lz + _r * math.sin(dth * j)))
    layer = np.array(layer)
    pos = np.concatenate([np.c_[layer, np.full(len(layer), r)]] * int(ny))
    return _radii(pos, min_radius, max_radius, seed, 14), [len(layer) - 1] * int(ny)


# --------------------------------------------------------------- importers
def read_ms3d_ascii(path_or_text):
    """vpolygon::_loadMS3DASCII: (vertices (V, 3) in m, triangles (T, 3) vertex
    indices) of the first mesh of a MilkShape 3D ASCII file."""
    text = _text(path_or_text)
    tok = text.split()
    k = 10                                   # '//' 'MilkShape' '3D' 'ASCII' Frames: n Frame: n Meshes: n
    k += 1                                   # "mesh name"
    k += 2                                   # flags, material index
    nvert = int(tok[k]); k += 1
    v = np.array(tok[k:k + 7 * nvert], float).reshape(nvert, 7)
    k += 7 * nvert
    verts = 0.001 * v[:, 1:4]
    nnorm = int(tok[k]); k += 1 + 3 * nnorm
    ntri = int(tok[k]); k += 1
    t = np.array(tok[k:k + 8 * ntri], int).reshape(ntri, 8)
    return verts, t[:, 1:4].astype(np.int64)


def read_stl_ascii(path_or_text):
    """vpolygon::_loadSTLASCII: (T, 3, 3) triangle vertices in m (mm x 0.001)."""
    tok = _text(path_or_text).split()
    vals = [tok[i + 1:i + 4] for i, t in enumerate(tok) if t == "vertex"]
    if len(vals) % 3:
        raise ValueError("STL: vertex count is not a multiple of 3")
    return 0.001 * np.array(vals, float).reshape(-1, 3, 3)


def _tri_rad(p, q, r):
    """Circumradius as the reference measures it: distance from the first
    vertex to calculate_center_of_triangle (source/vpolygon.cpp:433-434)."""
    c = scenes.circumcenter(p, q, r)
    return float(np.linalg.norm(c - np.asarray(p)))


def _split_one(t, to):
    """vpolygon::_splitTriangle (source/vpolygon.cpp:312-424): rounds over the
    pieces still too large; a piece whose circumradius exceeds `to` is cut at
    the midpoint of its longest edge (strict comparisons as the reference) into
    two pieces with the reference's vertex orders, the others are final, in the
    order they are reached (level by level)."""
    done, pending = [], [t]
    while pending:
        nxt = []
        for p, q, r in pending:
            if _tri_rad(p, q, r) > to:
                s_pq, s_qr, s_pr = (np.linalg.norm(q - p), np.linalg.norm(r - q),
                                    np.linalg.norm(r - p))
                if s_pq > s_qr:
                    tid = 3 if s_pq > s_pr else 2
                else:
                    tid = 1 if s_qr > s_pr else 2
                if tid == 1:
                    m = 0.5 * (r + q)
                    nxt += [(p, q, m), (p, m, r)]
                elif tid == 2:
                    m = 0.5 * (r + p)
                    nxt += [(q, r, m), (q, m, p)]
                else:
                    m = 0.5 * (q + p)
                    nxt += [(r, p, m), (r, m, q)]
            else:
                done.append((p, q, r))
        pending = nxt
    return done


def split_triangles(tris, to):
    """vpolygon::splitTriangle (source/vpolygon.cpp:426-470), the .xdm tool
    `poly_refinement`: every triangle whose circumradius exceeds `to` is
    replaced by its refinement (longest-edge bisection until each piece's
    circumradius is at most `to`), in place in the triangle order.  Orientation
    and area are kept.  tris: (T, 3, 3); returns (T', 3, 3)."""
    to = float(to)
    if not to > 0:
        raise ValueError("refinement size must be > 0")
    out = []
    for P, Q, R in np.asarray(tris, dtype=np.float64):
        if _tri_rad(P, Q, R) > to:
            out += _split_one((P, Q, R), to)
        else:
            out.append((P, Q, R))
    return np.array(out, dtype=np.float64).reshape(-1, 3, 3)


def _text(path_or_text):
    if isinstance(path_or_text, (str, os.PathLike)) and os.path.exists(path_or_text):
        with open(path_or_text, "r", encoding="utf-8", errors="replace") as f:
            return f.read()
    return str(path_or_text)


def ms3d_triangles(verts, tris):
    """(T, 3, 3) vertices of the indexed triangles and the object origin of
    polygonObject::_fromMS3DASCII: the mean of the circumcentres."""
    P = verts[tris]
    cc = scenes.circumcenter(P[:, 0], P[:, 1], P[:, 2])
    return P, cc.mean(0)


# -------------------------------------------------------------------- .xdm
@dataclass
class ParticleGroup:
    shape: str
    name: str
    material_type: int
    params: dict
    material: tuple          # youngs, density, poisson, shear


@dataclass
class GeometryObject:
    shape: str
    name: str
    material_type: int
    use: int
    params: dict
    material: tuple          # youngs, poisson, density, shear (the order geometryObjects writes)


@dataclass
class ContactPair:
    name: str
    method: int
    object0: str
    object1: str
    restitution: float
    ratio: float
    friction: float
    cohesion: float


@dataclass
class Model:
    name: str = ""
    gravity: tuple = (0.0, -9.80665, 0.0)
    unit: int = 0
    single: bool = False
    integrator: int = 0
    sort: int = 0
    particles: list = field(default_factory=list)
    objects: list = field(default_factory=list)
    contacts: list = field(default_factory=list)
    has_mbd: bool = False
    base_dir: str = "."

    def _generate(self, seed):
        """Positions of every generator, and the layer sizes that the plane and
        circle generators append to particleManager::np_group (:362, :477)."""
        out, groups = [], []
        for k, g in enumerate(self.particles):
            p = g.params
            if g.shape == "cube":
                pos = cube_particles(p["DIMENSION"], p["STARTPOINT"], p["RADIUS"][0], seed + k)
                if p["RADIUS"][0] != p["RADIUS"][1]:
                    raise ValueError("the cube generator is monodisperse")
            elif g.shape == "plane":
                dx, ny, dz = p["DIMENSION"]
                pos, gs = plane_particles(dx, int(ny), dz, p["STARTPOINT"], *p["RADIUS"], seed + k)
                groups += gs
            else:
                pos, gs = circle_particles(p["DIAMETER"], int(p["NUM_HEIGHT"]), p["STARTPOINT"],
                                           *p["RADIUS"], seed + k)
                groups += gs
            out.append(pos)
        return (np.concatenate(out) if out else np.zeros((0, 4))), groups

    def particle_positions(self, seed=0):
        return self._generate(seed)[0]

    def release(self, dt, seed=0):
        """Real-time creation as the generators set it up (particleManager.cpp:392-397
        and :588-594; `MAKING realtime np per_np time_per_group one_by_one file`) and
        dem_simulation::initialize starts it (dem_simulation.cpp:210-222): the flags
        and rates of the last plane or circle generator apply, and groups release
        the layer sizes of every plane and circle generator in creation order.
        Returns the Scene.release tuple, or None."""
        last = [g for g in self.particles if g.shape in ("plane", "circle")]
        if not last or not int(last[-1].params["MAKING"][0]):
            return None
        g = last[-1]
        mk = g.params["MAKING"]
        if int(mk[4]):
            if g.shape == "plane":
                # CreatePlaneParticle sets per_np = np / ndim.y with ndim.y = 0
                # (particleManager.cpp:332, :395): the reference divides by zero
                raise ValueError("one-by-one release of the plane generator is undefined "
                                 "in the reference (its rate divides by zero)")
            every = scenes.release_period(dt, per_second=int(mk[2]))
            return (every, 1) if every else None
        every = scenes.release_period(dt, time_per_group=float(mk[3]))
        return (every, self._generate(seed)[1]) if every else None

    def to_scene(self, dt, nsteps, cell_size=0.0, grid_size=(128, 128, 128),
                 world_origin=(-1.0, -1.0, -1.0), seed=0, rolling=True):
        """A Scene with the generated particles, planes (plane objects and the six
        planes of each cube, in object order) and polygon objects (their meshes
        read relative to the model file).  Contact parameters: the
        particle-particle pair of CONTACT_ELEMENTS_DATA (ratio = DHS ks/kn);
        walls use the first particle-wall pair (the simulation takes one wall
        pair per scene)."""
        pos, _ = self._generate(seed)
        if not self.particles:
            raise ValueError("the model creates no particles")
        y, rho, nu, g = self.particles[0].material
        mat = scenes.MaterialSpec(density=rho, youngs=y, poisson=nu)
        pp = [c for c in self.contacts if c.object0 == "particles" and c.object1 == "particles"]
        pw = [c for c in self.contacts if (c.object0 == "particles") != (c.object1 == "particles")]
        c0 = (pp or pw or [ContactPair("", 1, "", "", 0.8, 0.8, 0.3, 0.0)])[0]
        contact = scenes.ContactSpec(restitution=c0.restitution, friction=c0.friction,
                                     cohesion=c0.cohesion, stiffness_ratio=c0.ratio)
        planes, tri, tri_obj, objs = [], [], [], []
        for o in self.objects:
            p = o.params
            if o.shape == "plane":
                planes.append(scenes.plane4(p["POINT_A"], p["POINT_B"], p["POINT_C"], p["POINT_D"]))
            elif o.shape == "cube":
                planes += scenes.cube_planes(p["START_POINT"], p["END_POINT"])
            elif o.shape == "polygon":
                f = os.path.join(self.base_dir, p["FILE"])
                kind = int(p["IMPORT_SHAPE_TYPE"])   # import_shape_type: 1 MilkShape, 2 STL
                if kind == 1:
                    P, origin = ms3d_triangles(*read_ms3d_ascii(f))
                elif kind == 2:
                    P, origin = read_stl_ascii(f), tuple(p["CM"])
                else:
                    raise ValueError(f"unsupported mesh format {kind}")
                if p["TOOL"][0] == "poly_refinement" and p["TOOL"][2]:
                    # geometryObjects.cpp:88-89: refine before the object is made
                    P = split_triangles(P, float(p["TOOL"][2]))
                tri.append(P)
                tri_obj.append(np.full(len(P), len(objs), np.int32))
                objs.append(scenes.PolygonObject(origin=tuple(origin)))
        sc = scenes.Scene(name=self.name or "xdm", pos=np.ascontiguousarray(pos), dt=dt,
                          nsteps=nsteps, gravity=tuple(self.gravity), world_origin=world_origin,
                          cell_size=cell_size, grid_size=grid_size, material=mat, contact=contact,
                          planes=planes, rolling=rolling, release=self.release(dt, seed))
        if tri:
            sc.tri = np.concatenate(tri)
            sc.tri_object = np.concatenate(tri_obj)
            sc.objects = objs
        return sc


_PARTICLE_FIELDS = {
    "cube": [("NAME", 1), ("MATERIAL_TYPE", 1), ("DIMENSION", 3), ("STARTPOINT", 3),
             ("SPACE", 1), ("RADIUS", 2), ("MATERIAL", 4)],
    "plane": [("NAME", 1), ("MATERIAL_TYPE", 1), ("DIMENSION", 3), ("STARTPOINT", 3),
              ("DIRECTION", 3), ("SPACE", 1), ("RADIUS", 2), ("MAKING", 6), ("MATERIAL", 4)],
    "circle": [("NAME", 1), ("MATERIAL_TYPE", 1), ("DIAMETER", 1), ("NUM_HEIGHT", 1),
               ("STARTPOINT", 3), ("DIRECTION", 3), ("SPACE", 1), ("RADIUS", 2), ("MAKING", 6),
               ("MATERIAL", 4)],
}
_OBJECT_FIELDS = {
    "cube": [("NAME", 1), ("MATERIAL_TYPE", 1), ("GEOMETRY_USE", 1), ("START_POINT", 3),
             ("END_POINT", 3), ("MATERIAL", 4)],
    "plane": [("NAME", 1), ("MATERIAL_TYPE", 1), ("GEOMETRY_USE", 1), ("POINT_A", 3),
              ("POINT_B", 3), ("POINT_C", 3), ("POINT_D", 3), ("MATERIAL", 4)],
}


def _num(s):
    try:
        return float(s)
    except ValueError:
        return s


class _Tokens:
    """Whitespace tokens with line access (a polygon object's file name is the
    rest of its line, as geometryObjects::Open reads it with readLine)."""

    def __init__(self, text):
        self.lines = text.splitlines()
        self.li, self.buf = 0, []

    def _fill(self):
        while not self.buf and self.li < len(self.lines):
            self.buf = self.lines[self.li].split()
            self.li += 1
        return bool(self.buf)

    def next(self):
        if not self._fill():
            raise ValueError(".xdm: unexpected end of file")
        return self.buf.pop(0)

    def rest_of_line(self):
        out = " ".join(self.buf)
        self.buf = []
        return out

    def more(self):
        return self._fill()


def _fields(t, spec):
    out = {}
    for key, n in spec:
        tag = t.next()
        if tag != key:
            raise ValueError(f".xdm: expected {key}, found {tag}")
        vals = [_num(t.next()) for _ in range(n)]
        out[key] = vals[0] if n == 1 else vals
    return out


def parse_xdm(text, base_dir="."):
    """A Model from .xdm text (modelManager::OpenModel)."""
    t = _Tokens(text)
    m = Model(base_dir=base_dir)
    while t.more():
        ch = t.next()
        if ch == "MODEL_NAME":
            m.name = t.next()
            t.next()
            m.gravity = tuple(float(t.next()) for _ in range(3))
            t.next()
            m.unit = int(t.next())
            t.next()
            m.single = bool(int(t.next()))
        elif ch == "DEM_MODEL_DATA":
            t.next()
            m.integrator = int(t.next())
            t.next()
            m.sort = int(t.next())
        elif ch == "PARTICLES_DATA":
            while (ch := t.next()) != "END_DATA":
                if ch == "CREATE_SHAPE":
                    shape = t.next()
                    f = _fields(t, _PARTICLE_FIELDS[shape])
                    m.particles.append(ParticleGroup(shape, str(f.pop("NAME")),
                                                     int(f.pop("MATERIAL_TYPE")), f,
                                                     tuple(f.pop("MATERIAL"))))
        elif ch == "GEOMETRY_OBJECTS_DATA":
            while (ch := t.next()) != "END_DATA":
                if ch != "OBJECT":
                    continue
                shape = t.next()
                if shape in _OBJECT_FIELDS:
                    f = _fields(t, _OBJECT_FIELDS[shape])
                else:   # polygon (the log of geometryObjects::makePolygonObject, :194-212)
                    f = {"NAME": (t.next(), t.next())[1]}
                    t.next()
                    f["FILE"] = t.rest_of_line()
                    f.update(_fields(t, [("CM", 3), ("MATERIAL_TYPE", 1), ("GEOMETRY_USE", 1),
                                         ("IMPORT_SHAPE_TYPE", 1), ("MATERIAL", 4)]))
                    tool = (t.next(), t.next(), t.next(), _num(t.next()))
                    f["TOOL"] = (tool[1], tool[2], tool[3])
                m.objects.append(GeometryObject(shape, str(f.pop("NAME")),
                                                int(f.pop("MATERIAL_TYPE")),
                                                int(f.pop("GEOMETRY_USE")), f,
                                                tuple(f.pop("MATERIAL"))))
        elif ch == "CONTACT_ELEMENTS_DATA":
            ch = t.next()
            while ch != "END_DATA":
                f = {"NAME": t.next()}
                f.update(_fields(t, [("METHOD", 1), ("OBJECT0", 1), ("OBJECT1", 1),
                                     ("RESTITUTION", 1), ("RATIO", 1), ("FRICTION", 1),
                                     ("COHESION", 1)]))
                m.contacts.append(ContactPair(f["NAME"], int(f["METHOD"]), str(f["OBJECT0"]),
                                              str(f["OBJECT1"]), f["RESTITUTION"], f["RATIO"],
                                              f["FRICTION"], f["COHESION"]))
                ch = t.next()
        elif ch == "MULTIBODY_MODEL_DATA":
            m.has_mbd = True
            while t.more() and t.next() != "END_DATA":
                pass
    return m


def read_xdm(path):
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        return parse_xdm(f.read(), base_dir=os.path.dirname(os.path.abspath(path)))


def write_xdm(m: Model) -> str:
    """The .xdm text of a Model, in modelManager::SaveCurrentModel's layout
    (source/xcar/modelManager.cpp:30-53; section writers as listed above)."""
    def fmt(v):
        return " ".join(repr(float(x)) if isinstance(x, float) else str(x)
                        for x in (v if isinstance(v, (list, tuple)) else [v]))
    out = [f"MODEL_NAME {m.name}", f"GRAVITY {fmt(list(m.gravity))}", f"UNIT {m.unit}",
           f"SINGLE {int(m.single)}", "", "DEM_MODEL_DATA", f"INTEGRATOR {m.integrator}",
           f"SORT {m.sort}", "", "PARTICLES_DATA"]
    for g in m.particles:
        out.append(f"CREATE_SHAPE {g.shape}")
        for key, _ in _PARTICLE_FIELDS[g.shape]:
            v = (g.name if key == "NAME" else g.material_type if key == "MATERIAL_TYPE"
                 else list(g.material) if key == "MATERIAL" else g.params[key])
            out.append(f"{key} {fmt(v)}")
    out += ["END_DATA", "", "GEOMETRY_OBJECTS_DATA"]
    for o in m.objects:
        out.append(f"OBJECT {o.shape}")
        if o.shape == "polygon":
            p = o.params
            out += [f"NAME {o.name}", f"FILE {p['FILE']}", f"CM {fmt(p['CM'])}",
                    f"MATERIAL_TYPE {o.material_type}", f"GEOMETRY_USE {o.use}",
                    f"IMPORT_SHAPE_TYPE {int(p['IMPORT_SHAPE_TYPE'])}",
                    f"MATERIAL {fmt(list(o.material))}",
                    f"TOOL {p['TOOL'][0]} {p['TOOL'][1]} {fmt(p['TOOL'][2])}"]
            continue
        for key, _ in _OBJECT_FIELDS[o.shape]:
            v = (o.name if key == "NAME" else o.material_type if key == "MATERIAL_TYPE"
                 else o.use if key == "GEOMETRY_USE" else list(o.material) if key == "MATERIAL"
                 else o.params[key])
            out.append(f"{key} {fmt(v)}")
    out += ["END_DATA", "", "CONTACT_ELEMENTS_DATA"]
    for c in m.contacts:
        out += [f"NAME {c.name}", f"METHOD {c.method}", f"OBJECT0 {c.object0}",
                f"OBJECT1 {c.object1}", f"RESTITUTION {fmt(c.restitution)}",
                f"RATIO {fmt(c.ratio)}", f"FRICTION {fmt(c.friction)}",
                f"COHESION {fmt(c.cohesion)}"]
    out.append("END_DATA")
    return "\n".join(out) + "\n"
