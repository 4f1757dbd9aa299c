"""Model input (CPU): particle generators against their closed forms, the mesh
importers on the reference's own mesh file and on written text, and the .xdm
parser on text in the reference's save layout (modelManager::SaveCurrentModel
and the section writers: particleManager, geometryObjects, contactManager)."""
import math
import os

import numpy as np
import pytest

from metakonga_amd import model_io as mio
from metakonga_amd import scenes

REF_BOX = "/root/reference/resource/box.txt"


@pytest.mark.parametrize("dim,r", [((0.1, 0.1, 0.1), 5e-3), ((0.05, 0.12, 0.031), 2e-3),
                                   ((0.009, 0.1, 0.1), 2e-3)])
def test_cube_generator(dim, r):
    n = [max(int(L / (2 * r)) - 1, 0) for L in dim]
    pos = mio.cube_particles(dim, (0.1, -0.2, 0.3), r, seed=5)
    assert len(pos) == mio.num_cube_particles(*dim, r, r) == n[0] * n[1] * n[2]
    if not len(pos):
        return
    lo = np.array([0.1, -0.2, 0.3])
    space = (np.array(dim) - 2 * r * np.array(n)) / (np.array(n) + 1)
    # lattice node k of an axis at start + r + space + k (2r + space), jitter r * space * [0, 1]
    k = np.rint((pos[:, :3] - lo - r - space) / (2 * r + space)).astype(int)
    off = pos[:, :3] - (lo + r + space + k * (2 * r + space))
    assert (off >= -1e-15).all() and (off <= r * space * (1 + 1e-12) + 1e-15).all()
    assert (k.min(0) == 0).all() and (k.max(0) == np.array(n) - 1).all()
    # z outer, x inner
    assert np.array_equal(k[: n[0], 0], np.arange(n[0])) and (k[: n[0], 1:] == 0).all()
    # inside the box, no overlap
    assert (pos[:, :3] - r >= lo - 1e-15).all() and (pos[:, :3] + r <= lo + dim + 1e-12).all()
    with pytest.raises(ValueError):
        mio.num_cube_particles(*dim, r, 2 * r)


def test_circle_generator_rings():
    r, dia = 2e-3, 0.05
    pos, groups = mio.circle_particles(dia, 2, (0.01, 0.2, -0.03), r, r)
    assert len(pos) == mio.num_circle_particles(dia, 2, r, r)
    nr = int(0.5 * dia / (2 * r)) - 1
    space = (0.5 * dia - (2 * r * nr + r)) / (nr + 1)
    # dth = (2r + s) / (i (2r + s)) = 1 / i: floor(2 pi i) particles on ring i
    per_ring = [int(2 * math.pi * i) for i in range(1, nr + 1)]
    assert groups == [sum(per_ring)] * 2 and len(pos) == 2 * (1 + sum(per_ring))
    layer = pos[: 1 + sum(per_ring)]
    d = np.hypot(layer[:, 0] - 0.01, layer[:, 2] + 0.03)
    start = 1
    for i, m in enumerate(per_ring, 1):
        ring = d[start:start + m]
        assert np.allclose(ring, i * (2 * r + space), rtol=1e-13)
        start += m
    assert d[0] == 0 and (pos[:, 1] == 0.2).all() and (pos[:, 3] == r).all()


def test_plane_generator_and_radii():
    pos, groups = mio.plane_particles(0.1, 3, 0.06, (0.0, 0.5, 0.0), 2e-3, 4e-3, seed=9)
    n = (int(0.1 / 8e-3) - 1) * (int(0.06 / 8e-3) - 1)
    assert groups == [n] * 3 and len(pos) == mio.num_plane_particles(0.1, 3, 0.06, 2e-3, 4e-3)
    assert (pos[:, 1] == 0.5).all()
    assert pos[:, 3].min() >= 2e-3 and pos[:, 3].max() <= 4e-3 and pos[:, 3].std() > 1e-4
    mono, _ = mio.plane_particles(0.1, 1, 0.06, (0.0, 0.5, 0.0), 4e-3, 4e-3, seed=9)
    assert (mono[:, 3] == 4e-3).all()


MS3D = """// MilkShape 3D ASCII

Frames: 30
Frame: 1

Meshes: 1
"tri" 0 -1
4
0 0.0 0.0 0.0 0.0 0.0 -1
0 10.0 0.0 0.0 0.0 0.0 -1
0 0.0 10.0 0.0 0.0 0.0 -1
0 0.0 0.0 10.0 0.0 0.0 -1
1
0.0 0.0 1.0
2
0 0 1 2 0 0 0 1
0 0 2 3 0 0 0 1
"""

STL = """solid t
facet normal 0 0 1
  outer loop
    vertex 0 0 0
    vertex 20 0 0
    vertex 0 20 0
  endloop
endfacet
facet normal 1 0 0
  outer loop
    vertex 0 0 0
    vertex 0 20 0
    vertex 0 0 20
  endloop
endfacet
endsolid t
"""


def test_ms3d_and_stl_importers():
    v, t = mio.read_ms3d_ascii(MS3D)
    assert np.array_equal(v, 0.001 * np.array([[0, 0, 0], [10, 0, 0], [0, 10, 0], [0, 0, 10.0]]))
    assert np.array_equal(t, [[0, 1, 2], [0, 2, 3]])
    P, origin = mio.ms3d_triangles(v, t)
    # right triangles: circumcentres at the hypotenuse midpoints
    assert np.allclose(origin, 0.5 * (np.array([0.005, 0.005, 0]) + np.array([0, 0.005, 0.005])))
    tri = mio.read_stl_ascii(STL)
    assert tri.shape == (2, 3, 3) and np.array_equal(tri[1, 2], [0, 0, 0.02])


@pytest.mark.skipif(not os.path.exists(REF_BOX), reason="the reference mesh file is absent")
def test_reference_milkshape_box():
    v, t = mio.read_ms3d_ascii(REF_BOX)
    assert v.shape == (20, 3) and t.shape == (12, 3) and t.max() < 20
    assert np.allclose(v.min(0), [0.00025, 0.0, -0.01]) and np.allclose(v.max(0), [0.02, 0.02, 0.01])


# a model in the reference's save layout (the section writers' tags and order)
XDM = """MODEL_NAME hopper
GRAVITY 0 -9.80665 0
UNIT 0
SINGLE 0

DEM_MODEL_DATA
INTEGRATOR 0
SORT 0

PARTICLES_DATA
CREATE_SHAPE cube
NAME particles
MATERIAL_TYPE 0
DIMENSION 0.05 0.05 0.05
STARTPOINT 0.01 0.01 0.01
SPACE 0
RADIUS 0.005 0.005
MATERIAL 10000000 2500 0.3 3846153.8
END_DATA

GEOMETRY_OBJECTS_DATA
OBJECT cube
NAME box
MATERIAL_TYPE 0
GEOMETRY_USE 0
START_POINT 0 0 0
END_POINT 0.1 0.1 0.1
MATERIAL 10000000 0.3 2500 3846153.8
OBJECT plane
NAME ramp
MATERIAL_TYPE 0
GEOMETRY_USE 0
POINT_A 0 0.02 0
POINT_B 0.1 0.02 0
POINT_C 0.1 0.01 0.1
POINT_D 0 0.01 0.1
MATERIAL 10000000 0.3 2500 3846153.8
OBJECT polygon
NAME fin
FILE fin mesh.stl
CM 0.01 0.02 0.03
MATERIAL_TYPE 0
GEOMETRY_USE 0
IMPORT_SHAPE_TYPE 2
MATERIAL 10000000 0.3 2500 3846153.8
TOOL none none 0
END_DATA

CONTACT_ELEMENTS_DATA
NAME pp
METHOD 1
OBJECT0 particles
OBJECT1 particles
RESTITUTION 0.7
RATIO 0.8
FRICTION 0.4
COHESION 0
NAME pw
METHOD 1
OBJECT0 particles
OBJECT1 box
RESTITUTION 0.6
RATIO 0.8
FRICTION 0.5
COHESION 0
END_DATA
"""


def test_xdm_parse_and_scene(tmp_path):
    (tmp_path / "fin mesh.stl").write_text(STL)
    (tmp_path / "m.xdm").write_text(XDM)
    m = mio.read_xdm(tmp_path / "m.xdm")
    assert m.name == "hopper" and m.gravity == (0.0, -9.80665, 0.0) and not m.single
    assert [g.shape for g in m.particles] == ["cube"]
    assert [o.shape for o in m.objects] == ["cube", "plane", "polygon"]
    assert m.objects[2].params["FILE"] == "fin mesh.stl" and m.objects[2].params["CM"] == [0.01, 0.02, 0.03]
    assert [c.name for c in m.contacts] == ["pp", "pw"] and m.contacts[1].friction == 0.5
    sc = m.to_scene(dt=1e-5, nsteps=10, grid_size=(32, 32, 32), world_origin=(-0.01, -0.01, -0.01))
    assert sc.np == 4 ** 3 and len(sc.planes) == 6 + 1
    assert sc.tri.shape == (2, 3, 3) and sc.objects[0].origin == (0.01, 0.02, 0.03)
    assert sc.contact.restitution == 0.7 and sc.contact.friction == 0.4
    assert sc.material.density == 2500 and sc.material.youngs == 1e7
    # the written model reads back the same
    again = mio.parse_xdm(mio.write_xdm(m), base_dir=str(tmp_path))
    assert again == m


def test_xdm_rejects_wrong_tags():
    bad = XDM.replace("START_POINT", "BEGIN_POINT")
    with pytest.raises(ValueError):
        mio.parse_xdm(bad)


def _area(t):
    return 0.5 * np.linalg.norm(np.cross(t[..., 1, :] - t[..., 0, :], t[..., 2, :] - t[..., 0, :]), axis=-1)


def test_split_right_triangle_exact():
    """vpolygon::splitTriangle on a right triangle (legs 2, circumradius sqrt 2)
    at size 1: the hypotenuse QR is the longest edge, cut at its midpoint into
    (P, Q, M) and (P, M, R), each of circumradius exactly 1 (final)."""
    P, Q, R = np.array([0.0, 0, 0]), np.array([2.0, 0, 0]), np.array([0.0, 2, 0])
    out = mio.split_triangles(np.array([[P, Q, R]]), 1.0)
    M = np.array([1.0, 1.0, 0.0])
    assert np.array_equal(out, np.array([[P, Q, M], [P, M, R]]))
    # already small enough: unchanged
    assert np.array_equal(mio.split_triangles(np.array([[P, Q, R]]), 1.5), np.array([[P, Q, R]]))


def test_split_level_order():
    """Pieces are final in the order the reference's rounds reach them: the
    first cut of this obtuse triangle leaves (P, M, R) small enough and (P, Q,
    M) not, so (P, M, R) comes before the pieces of (P, Q, M)."""
    P, Q, R = np.array([0.0, 0, 0]), np.array([4.0, 0, 0]), np.array([0.0, 1, 0])
    M = 0.5 * (R + Q)
    out = mio.split_triangles(np.array([[P, Q, R]]), 1.2)
    assert np.array_equal(out[0], np.array([P, M, R])) and len(out) == 3
    assert np.isclose(_area(out[1:]).sum(), _area(np.array([P, Q, M])), rtol=1e-14)


def test_split_properties():
    """Random triangles: every piece within the size (the reference's measure),
    pieces in the parent's plane with its orientation, areas adding up."""
    u = scenes.uniform(31, 60 * 9, 5).reshape(60, 3, 3) * 0.02
    rad = np.array([mio._tri_rad(*t) for t in u])
    to = 0.35 * rad.max()
    out = mio.split_triangles(u, to)
    assert len(out) > len(u)
    assert max(mio._tri_rad(*t) for t in out) <= to
    assert np.isclose(_area(out).sum(), _area(u).sum(), rtol=1e-12)
    # in place: the pieces of each parent follow each other, in its plane, with
    # its orientation and its area
    k = 0
    for t in u:
        pieces = mio.split_triangles(t[None], to)
        assert np.array_equal(out[k:k + len(pieces)], pieces)
        k += len(pieces)
        n = np.cross(t[1] - t[0], t[2] - t[0])
        n /= np.linalg.norm(n)
        pn = np.cross(pieces[:, 1] - pieces[:, 0], pieces[:, 2] - pieces[:, 0])
        assert (pn @ n > 0).all()
        assert np.abs((pieces - t[0]) @ n).max() < 1e-15
        assert np.isclose(_area(pieces).sum(), _area(t[None])[0], rtol=1e-12)
    assert k == len(out)


def test_xdm_poly_refinement(tmp_path):
    """The .xdm tool `poly_refinement name size` refines the mesh before the
    object is made (geometryObjects.cpp:88-89)."""
    (tmp_path / "fin mesh.stl").write_text(STL)
    (tmp_path / "m.xdm").write_text(XDM.replace("TOOL none none 0", "TOOL poly_refinement fin 0.004"))
    m = mio.read_xdm(tmp_path / "m.xdm")
    sc = m.to_scene(dt=1e-5, nsteps=10, grid_size=(32, 32, 32), world_origin=(-0.01, -0.01, -0.01))
    base = mio.read_stl_ascii(STL)
    assert sc.tri.shape[0] > len(base) and (sc.tri_object == 0).all()
    assert max(mio._tri_rad(*t) for t in sc.tri) <= 0.004
    assert np.isclose(_area(sc.tri).sum(), _area(base).sum(), rtol=1e-12)
