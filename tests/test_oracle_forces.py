"""Pins of the oracle's contact mechanics against closed forms (CPU only).

Force law: DHSModel include/mphysics_cuda_impl.cuh:379-401 (Hertz normal
kn*delta^1.5, linear normal damping, tangential min(ks|vt|dt + vs|vt|, mu|Fn|));
constants getConstant :214-256; plane detection :487-570; closest point on a
triangle :837-899 (Ericson 5.1.5); Verlet :85-134; reduction :776-835;
circumcentre source/numeric_utility.cpp:20-45.
"""
import math

import numpy as np
import pytest

from metakonga_amd import abi, scenes

MAT = scenes.MaterialSpec()
CON = scenes.ContactSpec()
CP = scenes.contact_property(MAT, MAT, CON)
G = (0.0, -9.80665, 0.0)


def prm(dt=1e-5, cs=0.01, wo=(-1, -1, -1), gs=(128, 128, 128), np_=0, rolling=True):
    d = abi.DeviceParameters()
    d.rollingCondition = rolling
    d.np = np_
    d.grid_size = abi.uint3(*gs)
    d.ncell = gs[0] * gs[1] * gs[2]
    d.dt = dt
    d.half2dt = 0.5 * dt * dt
    d.cell_size = cs
    d.gravity = abi.double3(*G)
    d.world_origin = abi.double3(*wo)
    return d


def estar(E=1e7, nu=0.3):
    # E* = E_i E_j / (E_i (1 - nu_j^2) + E_j (1 - nu_i^2)) for identical materials = E / (2 (1 - nu^2))
    return E / (2.0 * (1.0 - nu * nu))


def test_constants_appendix_c(orc):
    """Known answers for r = 5 mm, rho 2500, E 1e7, nu 0.3, e 0.8, sratio 0.8."""
    r = 5e-3
    m = 2500 * 4 / 3 * math.pi * r ** 3
    assert m == pytest.approx(1.308997e-3, rel=1e-6)
    pp = orc.force_constant(1, r, r, m, m, CP)
    assert pp.kn == pytest.approx(3.663004e5, rel=1e-6)
    assert pp.vn == pytest.approx(2.194041, rel=1e-6)
    assert pp.ks == pytest.approx(2.930403e5, rel=1e-6)
    assert pp.vs == pytest.approx(1.755233, rel=1e-6)
    pw = orc.force_constant(1, r, 0.0, m, 0.0, CP)
    assert pw.kn == pytest.approx(5.180269e5, rel=1e-6)
    assert pw.vn == pytest.approx(3.689922, rel=1e-6)
    assert pw.ks == pytest.approx(4.144216e5, rel=1e-6)
    assert pw.vs == pytest.approx(2.951938, rel=1e-6)
    # closed forms: kn = 4/3 E* sqrt(R*), vn = sqrt(4 M* kn / (1 + (pi/ln e)^2))
    R = r / 2
    M = m / 2
    kn = 4.0 / 3.0 * estar() * math.sqrt(R)
    beta = math.pi / math.log(0.8)
    assert pp.kn == pytest.approx(kn, rel=1e-14)
    assert pp.vn == pytest.approx(math.sqrt(4 * M * kn / (1 + beta ** 2)), rel=1e-14)
    # HMCM variant: ks = 8 G* sqrt(R*), G* = G/(2 (2 - nu))
    h = orc.force_constant(0, r, r, m, m, CP)
    Gs = MAT.shear / (2 * (2 - 0.3))
    assert h.ks == pytest.approx(8 * Gs * math.sqrt(R), rel=1e-14)


def _pair_state(d, vrel=(0, 0, 0), omega_i=(0, 0, 0), omega_j=(0, 0, 0), r=5e-3):
    pos = np.array([[0.0, 0.0, 0.0, r], [2 * r - d, 0.0, 0.0, r]])
    vel = np.array([[0.0, 0.0, 0.0], list(vrel)], dtype=np.float64)
    om = np.array([list(omega_i), list(omega_j)], dtype=np.float64)
    return pos, vel, om


def _p2p(orc, pos, vel, om, rolling=True, tcm=1):
    n = pos.shape[0]
    p = prm(np_=n, rolling=rolling)
    h, idx = orc.hash_and_index(p, pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    m, I = scenes.mass_inertia(pos[:, 3], MAT.density)
    f = np.zeros((n, 3))
    t = np.zeros((n, 3))
    orc.p2p(p, tcm, pos, vel, om, f, t, m, srt, cs, ce, CP)
    return f, t, m


def test_head_on_hertz(orc):
    r = 5e-3
    for d in (1e-6, 1.0709e-5, 4.6335e-5, 2e-4):
        pos, vel, om = _pair_state(d)
        f, t, m = _p2p(orc, pos, vel, om)
        kn = 4.0 / 3.0 * estar() * math.sqrt(r / 2)
        # on particle 0 the force points away from particle 1 (-x)
        assert f[0, 0] == pytest.approx(-kn * d ** 1.5, rel=1e-12)
        assert f[1, 0] == pytest.approx(kn * d ** 1.5, rel=1e-12)
        assert abs(f[0, 1]) == 0 and abs(f[0, 2]) == 0
        assert np.all(t == 0)


def test_normal_damping(orc):
    r = 5e-3
    d = 2e-5
    vn_rel = -0.1  # particle 1 approaching particle 0
    pos, vel, om = _pair_state(d, vrel=(vn_rel, 0, 0))
    f, t, m = _p2p(orc, pos, vel, om)
    M = m[0] / 2
    kn = 4.0 / 3.0 * estar() * math.sqrt(r / 2)
    beta = math.pi / math.log(0.8)
    vn = math.sqrt(4 * M * kn / (1 + beta ** 2))
    # dv = v_j - v_i = (-0.1,0,0); unit = +x; Fn on i = (-kn d^1.5 + vn * (-0.1)) * (+x)
    assert f[0, 0] == pytest.approx(-kn * d ** 1.5 + vn * vn_rel, rel=1e-12)


@pytest.mark.parametrize("vt", [1e-4, 0.5])
def test_tangential_closed_form(orc, vt):
    """Sliding contact: Ft = min(ks vt dt + vs vt, mu |Fn|) along the relative
    tangential velocity; torque (r n) x Ft (GPU lever arm)."""
    r = 5e-3
    d = 3e-5
    pos, vel, om = _pair_state(d, vrel=(0, 0, vt))
    f, t, m = _p2p(orc, pos, vel, om)
    M = m[0] / 2
    kn = 4.0 / 3.0 * estar() * math.sqrt(r / 2)
    beta = math.pi / math.log(0.8)
    vn = math.sqrt(4 * M * kn / (1 + beta ** 2))
    ks, vs = 0.8 * kn, 0.8 * vn
    Fn = kn * d ** 1.5
    ft = min(ks * vt * 1e-5 + vs * vt, 0.3 * Fn)
    assert f[0, 2] == pytest.approx(ft, rel=1e-12)
    # M = (r * unit) x Ft with unit=+x and Ft=+z: (r,0,0) x (0,0,ft) = (0, -r ft, 0)
    assert t[0, 1] == pytest.approx(-r * ft, rel=1e-12)
    if vt > 0.1:
        assert ft == pytest.approx(0.3 * Fn, rel=1e-15)  # saturated (Coulomb)


def test_rolling_switch(orc):
    """With rollingCondition false the spin terms drop out of dv (:455)."""
    pos, vel, om = _pair_state(2e-5, omega_i=(0, 0, 10.0))
    f_roll, _, _ = _p2p(orc, pos, vel, om, rolling=True)
    f_no, _, _ = _p2p(orc, pos, vel, om, rolling=False)
    assert f_no[0, 1] == 0.0
    assert f_roll[0, 1] != 0.0


def test_host_lever_arm(orc):
    """CPU path torque uses rcon = r - delta/2 (A1): ratio rcon / r to the GPU torque."""
    r = 5e-3
    d = 4e-5
    pos, vel, om = _pair_state(d, vrel=(0, 0, 1e-3))
    p = prm(np_=2)
    h, idx = orc.hash_and_index(p, pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    m, _ = scenes.mass_inertia(pos[:, 3], 2500.0)
    out = {}
    for sem in (orc.ORC_GPU, orc.ORC_HOST):
        f = np.zeros((2, 3))
        t = np.zeros((2, 3))
        orc.p2p(p, 1, pos, vel, om, f, t, m, srt, cs, ce, CP, semantics=sem)
        out[sem] = (f, t)
    assert np.allclose(out[0][0], out[1][0], rtol=1e-14, atol=0)
    assert out[1][1][0, 1] == pytest.approx(out[0][1][0, 1] * (r - 0.5 * d) / r, rel=1e-12)


def _floor_plane(L=1.0):
    return scenes.plane4((-L / 2, 0, -L / 2), (L / 2, 0, -L / 2), (L / 2, 0, L / 2), (-L / 2, 0, L / 2))


def test_sphere_on_plane(orc):
    """Face branch: F = kn_w delta^1.5 along the outward normal; no torque (the
    GPU kernel calls the force law with ir = 0, :609,614)."""
    r = 5e-3
    pl = _floor_plane()
    for d in (1e-6, 1.0709e-5):
        pos = np.array([[0.01, r - d, 0.02, r]])
        vel = np.zeros((1, 3))
        om = np.array([[0.0, 0.0, 3.0]])
        f = np.zeros((1, 3))
        t = np.zeros((1, 3))
        m, _ = scenes.mass_inertia(pos[:, 3], 2500.0)
        orc.plane(prm(np_=1), 1, pl, pos, vel, om, f, t, m, CP)
        knw = 4.0 / 3.0 * estar() * math.sqrt(r)
        assert f[0, 1] == pytest.approx(knw * d ** 1.5, rel=1e-12)
        assert np.all(t == 0)
    # static overlap under one particle's weight with the particle-particle
    # stiffness: delta = (m g / kn)^(2/3) = 1.0709e-5 m (known answer)
    mg = 2500 * 4 / 3 * math.pi * r ** 3 * 9.80665
    assert mg == pytest.approx(1.2837e-2, rel=1e-4)
    assert (mg / (4.0 / 3.0 * estar() * math.sqrt(r / 2))) ** (2 / 3) == pytest.approx(1.0709e-5, rel=1e-4)


def test_plane_corner_and_edge_branches(orc):
    r = 5e-3
    pl = _floor_plane(L=0.2)
    # corner xw=(-0.1,0,-0.1): a point beyond both edges, within r of the corner
    p = np.array([-0.1 - 0.002, 0.001, -0.1 - 0.002])
    dp = p - np.array([-0.1, 0, -0.1])
    wp = np.array([dp @ [1, 0, 0], dp @ [0, 0, 1], dp @ [0, -1, 0]])  # u1=x, u2=z, uw=u1 x u2=-y
    c, u = orc.plane_detect(pl, p, wp, r)
    h = np.linalg.norm(dp)
    assert c == pytest.approx(r - h, rel=1e-14)
    assert np.allclose(u, dp / h, rtol=1e-14)
    # edge along u1 at wp.y < 0: distance to the edge line xw->w2
    p = np.array([0.03, 0.002, -0.1 - 0.003])
    dp = p - np.array([-0.1, 0, -0.1])
    wp = np.array([dp[0], dp[2], -dp[1]])
    c, u = orc.plane_detect(pl, p, wp, r)
    assert c == pytest.approx(r - math.hypot(0.002, 0.003), rel=1e-12)
    # outside every branch
    p = np.array([0.5, 0.5, 0.5])
    dp = p - np.array([-0.1, 0, -0.1])
    c, _ = orc.plane_detect(pl, p, np.array([dp[0], dp[2], -dp[1]]), r)
    assert c == -1.0


def test_cube_planes_have_zero_corners(orc):
    """cube::define uses the 3-point plane form: w2 = w3 = w4 = 0, so the
    corner branch of w2 measures the distance to the origin (reproduced)."""
    planes = scenes.cube_planes((1.0, 1.0, 1.0), (1.2, 1.2, 1.2))
    pl = planes[3]  # z = 1 plane, xw = min, u1 = x, u2 = y
    assert pl.w2.x == 0 and pl.w3.y == 0 and pl.w4.z == 0
    r = 5e-3
    p = np.array([1.2 + 0.001, 1.0 - 0.001, 1.0])
    dp = p - np.array([1.0, 1.0, 1.0])
    wp = np.array([dp[0], dp[1], dp[2]])
    c, u = orc.plane_detect(pl, p, wp, r)
    # branch 2 (wp.x > l1, wp.y < 0): uses w2 = (0,0,0) -> r - |p| (no contact)
    assert c == -1.0 or c == pytest.approx(r - np.linalg.norm(p), rel=1e-12)


def test_host_plane_assignment(orc):
    """CPU path: a lone plane assigns F (A2), so of two planes in contact only
    the last one counts; the GPU kernel adds both."""
    r = 5e-3
    sc = scenes.Scene(name="t", pos=np.array([[0.0, r - 1e-5, 0.0, r]]), dt=1e-5, nsteps=1,
                      world_origin=(-1, -1, -1), cell_size=0.01,
                      planes=[_floor_plane(), _floor_plane()])
    out = {}
    for sem in (orc.ORC_GPU, orc.ORC_HOST):
        w = orc.World(sc, semantics=sem)
        w.force[:] = 0
        w.step(1)
        out[sem] = w.acc.copy()
    m = sc.mass_inertia()[0][0]
    knw = 4.0 / 3.0 * estar() * math.sqrt(r)
    # GPU: 2 planes; host: 1 plane (the velocity is zero at the first step)
    assert out[orc.ORC_GPU][0, 1] * m == pytest.approx(2 * knw * (1e-5) ** 1.5, rel=1e-9)
    assert out[orc.ORC_HOST][0, 1] * m == pytest.approx(knw * (1e-5) ** 1.5, rel=1e-9)


def _brute_closest(P, Q, R, p, n=400):
    a = np.linspace(0, 1, n)
    u, v = np.meshgrid(a, a, indexing="ij")
    keep = (u + v) <= 1
    u, v = u[keep], v[keep]
    pts = P + u[:, None] * (Q - P) + v[:, None] * (R - P)
    d = np.linalg.norm(pts - p, axis=1)
    return pts[d.argmin()], d.min()


def test_closest_point_triangle(orc):
    rng = scenes.uniform(77, 600, 1).reshape(-1, 3) * 2 - 1
    P, Q, R = np.array([0.0, 0, 0]), np.array([1.0, 0.1, 0]), np.array([0.2, 0.9, 0.1])
    tri = abi.PolygonInfo(0, abi._d3(P), abi._d3(Q), abi._d3(R))
    for k in range(rng.shape[0] // 2):
        p = rng[2 * k] * 1.5 + 0.3
        c = orc.closest_pt_triangle(tri, p)
        cb, db = _brute_closest(P, Q, R, p)
        assert np.linalg.norm(c - p) <= db + 1e-12
        assert np.linalg.norm(c - cb) < 5e-3


def test_verlet_parabola(orc):
    """Constant force m g: after n steps the GPU-form Verlet gives
    x = x0 + v0 t + g t^2 / 2 (exact in exact arithmetic)."""
    p = prm(dt=1e-3)
    pos = np.array([[0.1, 0.2, 0.3, 5e-3]])
    vel = np.array([[1.0, 2.0, -0.5]])
    acc = np.zeros((1, 3))
    om = np.zeros((1, 3))
    al = np.zeros((1, 3))
    m = np.array([2.0])
    I = np.array([1.0])
    f = m[:, None] * np.array(G)[None, :]
    t = np.zeros((1, 3))
    n = 200
    for _ in range(n):
        orc.vv_position(p, pos, vel, acc)
        orc.vv_velocity(p, vel, acc, om, al, f, t, m, I)
    T = n * 1e-3
    # first step uses a = 0 (acc starts at zero, dem_simulation.cpp:234): half-step kick missing
    exp_v = np.array([1.0, 2.0, -0.5]) + np.array(G) * T - np.array(G) * 0.5e-3
    assert np.allclose(vel[0], exp_v, rtol=1e-12, atol=1e-12)
    # x_n = x0 + v0 T + g (T^2 - T dt) / 2 (sum of the kicks with a_0 = 0)
    x_exact = (np.array([0.1, 0.2, 0.3]) + np.array([1.0, 2.0, -0.5]) * T
               + 0.5 * np.array(G) * (T * T - T * 1e-3))
    assert np.allclose(pos[0, :3], x_exact, rtol=0, atol=1e-12)


def test_reduction_exact(orc):
    for n in (1, 100, 511, 512, 513, 5000, 100003):
        a = (scenes.splitmix64(n, 3 * n, 1) % np.uint64(1000)).astype(np.float64).reshape(n, 3)
        s = orc.reduction_d3(a)
        assert np.array_equal(s, a.sum(axis=0))


def test_circumcenter_equidistant(orc):
    u = scenes.uniform(5, 90, 2).reshape(-1, 3, 3)
    for P, Q, R in u:
        c = orc.circumcenter(P, Q, R)
        d = [np.linalg.norm(c - x) for x in (P, Q, R)]
        assert max(d) - min(d) < 1e-9 * max(d)
        n = np.cross(Q - P, R - P)
        assert abs((c - P) @ n) < 1e-9 * np.linalg.norm(n) * max(d)
    # the numpy restatement used by the scene generator agrees
    c2 = scenes.circumcenter(u[:, 0], u[:, 1], u[:, 2])
    c1 = np.array([orc.circumcenter(P, Q, R) for P, Q, R in u])
    assert np.allclose(c1, c2, rtol=1e-12, atol=1e-14)


def test_cylinder_lateral(orc):
    """Solid cylinder, lateral region: cdist = rtop + r - dist (rounded to FP32)."""
    cy = abi.CylinderInfo()
    cy.len = 1.0
    cy.rbase = cy.rtop = 0.2
    cy.pbase = abi.double3(0, -0.5, 0)
    cy.ptop = abi.double3(0, 0.5, 0)
    cy.origin = abi.double3(0, 0, 0)
    cy.ep = abi.double4(1, 0, 0, 0)
    r = 5e-3
    p = np.array([0.2 + r - 1e-4, 0.1, 0.0, r])
    c, u, cp = orc.cylinder_detect(cy, p)
    assert c == float(np.float32(0.2 + r - (0.2 + r - 1e-4)))
    assert np.allclose(u, [-1, 0, 0])
    assert np.allclose(cp, [0.2, 0.1, 0.0], atol=1e-15)


def _cyl(axis_ep=(1.0, 0.0, 0.0, 0.0), L=1.0, R=0.2):
    """Solid cylinder of length L and radius R centred at the origin, its axis
    the local y axis turned by the Euler parameters (e0 first)."""
    e0, e1, e2, e3 = axis_ep
    ax = np.array([2 * (e1 * e2 - e0 * e3), 2 * (e0 * e0 + e2 * e2 - 0.5), 2 * (e2 * e3 + e0 * e1)])
    cy = abi.CylinderInfo()
    cy.len = L
    cy.rbase = cy.rtop = R
    cy.pbase = abi._d3(-0.5 * L * ax)
    cy.ptop = abi._d3(0.5 * L * ax)
    cy.origin = abi.double3(0, 0, 0)
    cy.ep = abi.double4(*axis_ep)
    return cy, ax


@pytest.mark.parametrize("theta", [0.0, 0.7])
def test_cylinder_cap_and_rim(orc, theta):
    """Cap region (beyond an end face, within the radius): cdist = r - h for a
    particle h above the face, normal along the axis towards the centre.  Rim
    region (beyond the face and outside the radius): the distance to the rim
    circle, sqrt((rho - R)^2 + h^2), normal towards the rim point.  Both ends,
    for an axis along y and one turned by theta about x (the rim is found in
    the cylinder's own frame).  Values rounded to FP32 as the reference's
    float return."""
    ep = (math.cos(theta / 2), math.sin(theta / 2), 0.0, 0.0)
    cy, ax = _cyl(ep)
    L, R, r = 1.0, 0.2, 5e-3
    # a radial direction perpendicular to the axis
    e = np.cross(ax, [1.0, 0.0, 0.0])
    e = e / np.linalg.norm(e) if np.linalg.norm(e) > 0 else np.array([0.0, 0.0, 1.0])
    for side in (1.0, -1.0):
        face = side * 0.5 * L * ax
        # cap: lateral offset 0.1 < R, height h above the face
        h = r - 1e-4
        p = face + side * h * ax + 0.1 * e
        c, u, cp = orc.cylinder_detect(cy, np.r_[p, r])
        assert c == float(np.float32(r - h)), (side, c)
        assert np.allclose(u, -side * ax, atol=1e-12)
        # rim: radial offset R + a, height b above the face, distance < r
        a, b = 2e-3, 3e-3
        p = face + side * b * ax + (R + a) * e
        c, u, cp = orc.cylinder_detect(cy, np.r_[p, r])
        d = math.hypot(a, b)
        assert abs(c - (r - d)) <= 2.0 ** -23 * (r - d), (side, c, r - d)   # FP32 return
        rim = face + R * e
        assert np.allclose(cp, rim, atol=1e-12)
        assert np.allclose(u, (rim - p) / d, atol=1e-9)
        # farther than r from the rim: no contact
        p = face + side * 6e-3 * ax + (R + 6e-3) * e
        assert orc.cylinder_detect(cy, np.r_[p, r])[0] < 0


def _tri_scene(pos):
    """A static 6 mm triangle in the plane y = 0 (face normal -(Q-P)x(R-P) = -y),
    object origin (0.01, 0, 0.02)."""
    P, Q, R = (-3e-3, 0.0, -3e-3), (-3e-3, 0.0, 3e-3), (3e-3, 0.0, 0.0)
    return scenes.Scene(name="tri", pos=np.ascontiguousarray(pos), dt=1e-5, nsteps=1,
                        world_origin=(-0.05, -0.05, -0.05), grid_size=(32, 32, 32),
                        tri=np.array([[P, Q, R]]),
                        objects=[scenes.PolygonObject(origin=(0.01, 0.0, 0.02))])


def test_polygon_reaction_closed_form(orc):
    """Particles at rest against a static triangle: each contact's force is
    Hertz kn delta^1.5 along the face normal (no damping, no tangential part);
    the object's reaction is minus the sum of the particle forces, its moment
    about the object origin -(cp - origin) x F summed (cp the foot of the
    perpendicular).  One particle, then two (the sums add up)."""
    r = 2e-3
    pts = [(0.5e-3, r - 1e-5, -0.5e-3), (-1.5e-3, r - 3e-5, 1.0e-3)]
    origin = np.array([0.01, 0.0, 0.02])
    for npart in (1, 2):
        pos = np.array([[x, y, z, r] for x, y, z in pts[:npart]])
        sc = _tri_scene(pos)
        w = orc.World(sc)
        w.step(1)      # from rest: the contact positions are the initial ones
        F = np.zeros(3)
        Mo = np.zeros(3)
        for k in range(npart):
            c = orc.force_constant(1, r, 0.0, w.mass[k], 0.0, sc.contact_pp())
            delta = r - pos[k, 1]
            f = c.kn * delta ** 1.5 * np.array([0.0, 1.0, 0.0])   # on the particle, +y
            cp = np.array([pos[k, 0], 0.0, pos[k, 2]])
            F += f
            Mo += np.cross(cp - origin, f)
        d = w.dpmi[0]
        got_f = np.array([d.force.x, d.force.y, d.force.z])
        got_m = np.array([d.moment.x, d.moment.y, d.moment.z])
        assert np.allclose(got_f, -F, rtol=1e-12, atol=0), (npart, got_f, -F)
        assert np.allclose(got_m, -Mo, rtol=1e-12, atol=1e-30), (npart, got_m, -Mo)
