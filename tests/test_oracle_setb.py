"""Pins of the oracle's restatement of the older entry points
(include/cu_dem_dec.cuh; oracle/xdem_oracle_b.cpp), CPU only.

Each check is a closed form, a definition or an independent restatement, not
the oracle's own formula retyped:
  * hash = the live-set oracle's hash; sort = numpy's stable argsort; cell
    ranges = run-length invariants; cend untouched where cstart is cleared;
  * launch bounds: a call for n items touches exactly the first
    min(T(n), np) items, T(n) = ceil(n / min(256, n)) * min(256, n);
  * Verlet = exact parabola under constant force (two half kicks);
  * p2p = head-on Hertz + damping closed form along the centre line, no
    tangential force component (the reference adds only the normal force),
    sliding moment (ri n) x min(ks |e| dt + vs |e|, mu |Fn|) e_hat, cohesion
    step (coh / mu) n for separated spheres with -coh/mu < Fn < 0;
  * cube = a sphere resting on the floor face feels kn d^1.5 upward, nothing
    outside the face rectangle;
  * shape = face contact kn d^1.5 along the triangle normal, edge contact
    along the line from the closest point, one contact per particle, body
    force = the sum of the particle forces;
  * polygon updates = rotation / cross-product closed forms on integers.
"""
import numpy as np
import pytest

from metakonga_amd import abi, scenes
from xdem_testutil import as4, dense_scene, pair_constants, setb_params


def _box(n_side=6, seed=3):
    sc, vel, om = dense_scene(n_side=n_side, seed=seed)
    return sc, vel, om


def test_hash_sort_ranges(orc):
    sc, vel, om = _box(8)
    n = sc.np
    p = setb_params(sc, n)
    B = orc.SetB(p)
    h, idx = B.hash_and_index(sc.pos)
    h_live, _ = orc.hash_and_index(sc.params(), sc.pos)
    order = np.argsort(h_live, kind="stable")
    assert np.array_equal(h, h_live[order])
    assert np.array_equal(idx, order.astype(np.uint32))
    nc = p.ncell
    cend0 = np.full(nc, 777, np.uint32)
    cs, ce, srt = B.reorder(h, idx, nc, cend=cend0)
    assert np.array_equal(srt, idx)
    present = np.unique(h)
    for c in present:
        run = np.nonzero(h == c)[0]
        assert cs[c] == run[0] and ce[c] == run[-1] + 1
    absent = np.setdiff1d(np.arange(nc), present)
    assert (cs[absent] == 0xFFFFFFFF).all()
    assert (ce[absent] == 777).all()  # cend is not cleared (cu_dem_def.cu:80)


@pytest.mark.parametrize("n_call,np_", [(300, 1000), (1000, 300), (100, 1000), (257, 257)])
def test_launch_bounds(orc, n_call, np_):
    sc, _, _ = _box(4)
    p = setb_params(sc, np_)
    B = orc.SetB(p)
    N = 1200
    pos = np.zeros((N, 4))
    vel = as4(np.ones((N, 3)), 0.0)
    acc = as4(np.zeros((N, 3)), 1.0)
    B.vv_position(pos, vel, acc, n_call)
    t = min(256, n_call)
    threads = -(-n_call // t) * t
    touched = min(threads, np_)
    assert (pos[:touched, :3] == sc.dt).all()
    assert (pos[touched:, :3] == 0).all()


def test_verlet_parabola(orc):
    sc, _, _ = _box(4)
    n = 64
    p = setb_params(sc, n)
    B = orc.SetB(p)
    rng = np.random.default_rng(1)
    m = rng.uniform(1e-3, 2e-3, n)
    F = rng.normal(size=(n, 3))
    v0 = rng.normal(size=(n, 3))
    pos = as4(np.zeros((n, 3)), 5e-3)
    vel = as4(v0, 0.0)
    acc = as4(F / m[:, None], m)
    om = as4(np.zeros((n, 3)), 0.0)
    al = as4(np.zeros((n, 3)), 1e-8)
    mom = np.zeros((n, 3))
    K = 50
    for _ in range(K):
        B.vv_position(pos, vel, acc)
        B.vv_velocity(vel, acc, om, al, F, mom)
    t = K * sc.dt
    assert np.allclose(pos[:, :3], v0 * t + 0.5 * (F / m[:, None]) * t * t, rtol=1e-10, atol=1e-18)
    assert np.allclose(vel[:, :3], v0 + F / m[:, None] * t, rtol=1e-12)
    assert (acc[:, 3] == m).all() and (vel[:, 3] == 0).all() and (al[:, 3] == 1e-8).all()


def _pair(orc, gap, v_rel, coh=0.0, mu=0.3):
    """Two spheres on the x axis, sphere 1 at +x moving with v_rel."""
    r = float(scenes.as_f32_exact(5e-3))
    pos3 = np.array([[0.0, 0.0, 0.0], [2 * r + gap, 0.0, 0.0]])
    sc = scenes.Scene(name="pair", pos=np.c_[pos3, [r, r]], dt=1e-5, nsteps=1,
                      world_origin=(-0.08, -0.08, -0.08), cell_size=2 * r,
                      grid_size=(16, 16, 16), gravity=(0.0, -9.80665, 0.0))
    kn, vn, ks, vs = pair_constants(orc, sc)
    p = setb_params(sc, 2, kn=kn, vn=vn, ks=ks, vs=vs, mu=mu, cohesive=coh)
    B = orc.SetB(p)
    m, _ = sc.mass_inertia()
    pos = np.ascontiguousarray(sc.pos)
    vel = as4(np.array([[0.0, 0.0, 0.0], v_rel]), 0.0)
    acc = as4(np.zeros((2, 3)), m)
    om = as4(np.zeros((2, 3)), 0.0)
    h, idx = B.hash_and_index(pos)
    cs, ce, srt = B.reorder(h, idx, p.ncell)
    f = np.zeros((2, 3))
    mo = np.zeros((2, 3))
    B.p2p(pos, vel, acc, om, f, mo, srt, cs, ce)
    return dict(f=f, m=mo, mass=m, r=r, kn=kn, vn=vn, ks=ks, vs=vs, mu=mu, g=np.array(sc.gravity),
                dt=sc.dt)


def test_p2p_head_on_closed_form(orc):
    delta, vx = 2e-5, -0.01  # sphere 1 approaching
    q = _pair(orc, -delta, [vx, 0.0, 0.0])
    n = np.array([1.0, 0, 0])  # from 0 towards 1
    fn = -q["kn"] * delta ** 1.5 + q["vn"] * vx  # dv . n, dv = v1 - v0
    want0 = q["mass"][0] * q["g"] + fn * n
    assert np.allclose(q["f"][0], want0, rtol=1e-12, atol=1e-15)
    assert np.abs(q["m"]).max() == 0


def test_p2p_sliding_moment_and_no_tangential_force(orc):
    delta, vt = 2e-5, 1e-3
    q = _pair(orc, -delta, [0.0, 0.0, vt])
    fn = q["kn"] * delta ** 1.5
    ft = min(q["ks"] * vt * q["dt"] + q["vs"] * vt, q["mu"] * fn)
    # moment on 0: (r n) x (ft z) with n = +x -> -r ft y
    assert np.allclose(q["m"][0], [0.0, -q["r"] * ft, 0.0], rtol=1e-12, atol=1e-20)
    # the tangential force is not added to the force (cu_dem_impl.cuh:248)
    assert q["f"][0][2] == 0.0
    assert np.allclose(q["f"][0] - q["mass"][0] * q["g"], [-fn, 0, 0], rtol=1e-12)


def test_p2p_cohesion_step(orc):
    coh, mu = 5e-3, 0.3
    gap = 1e-6  # separated: Fn = -kn gap^1.5 in (-coh/mu, 0)
    a = _pair(orc, gap, [0.0, 0.0, 0.0], coh=coh, mu=mu)
    b = _pair(orc, gap, [0.0, 0.0, 0.0], coh=0.0, mu=mu)
    assert -coh / mu < -a["kn"] * gap ** 1.5 < 0
    assert np.allclose(a["f"][0] - b["f"][0], [coh / mu, 0, 0], rtol=1e-12)
    far = _pair(orc, 1e-3, [0.0, 0.0, 0.0], coh=coh, mu=mu)  # beyond the cohesion band
    assert np.array_equal(far["f"][0], b["f"][0])
    # separated and sliding: moment (r n) x ((Fn mu + coh) s_hat)
    vt = 1e-3
    c = _pair(orc, gap, [0.0, 0.0, vt], coh=coh, mu=mu)
    Fn = -c["kn"] * gap ** 1.5
    assert np.allclose(c["m"][0], [0.0, -c["r"] * (Fn * mu + coh), 0.0], rtol=1e-10)


def _floor_planes(L=0.2, h=0.0):
    """Six planes of an axis-aligned box [0, L]^3 (floor y = h first)."""
    def pl(xw, u1, u2, uw, l1, l2):
        p = abi.BPlaneInfo()
        p.l1, p.l2 = l1, l2
        p.u1, p.u2, p.uw, p.xw = (abi._d3(v) for v in (u1, u2, uw, xw))
        return p
    return [pl((0, h, 0), (1, 0, 0), (0, 0, 1), (0, 1, 0), L, L),
            pl((0, L, 0), (1, 0, 0), (0, 0, 1), (0, -1, 0), L, L),
            pl((0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0), L, L),
            pl((L, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), L, L),
            pl((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), L, L),
            pl((0, 0, L), (1, 0, 0), (0, 1, 0), (0, 0, -1), L, L)]


def test_cube_face_only(orc):
    r, delta = 5e-3, 3e-5
    pos = np.array([[0.1, r - delta, 0.1, r],      # resting on the floor
                    [0.25, r - delta, 0.1, r],     # beyond the floor rectangle
                    [0.1, 0.1, 0.1, r]])           # free
    sc = scenes.Scene(name="c", pos=pos, dt=1e-5, nsteps=1, cell_size=2 * r, grid_size=(64, 64, 64))
    p = setb_params(sc, 3)
    B = orc.SetB(p)
    planes = abi.struct_array(abi.BPlaneInfo, _floor_planes())
    kn, vn, ks, vs, mu = 5e5, 3.0, 4e5, 2.5, 0.3
    vel = as4(np.zeros((3, 3)), 0.0)
    om = as4(np.zeros((3, 3)), 0.0)
    f = np.zeros((3, 3))
    m = np.zeros((3, 3))
    B.cube(planes, (kn, vn, ks, vs, mu), np.ascontiguousarray(pos), vel, om, f, m)
    assert np.allclose(f[0], [0, kn * delta ** 1.5, 0], rtol=1e-12)
    assert (f[1] == 0).all() and (f[2] == 0).all() and (m == 0).all()


def _tri_mesh(y=0.0, n=6, h=4e-3, x0=0.05, z0=0.05):
    """A flat triangulated square in the plane y (normal +y), vertex grid
    n x n with pitch h; returns vertices, index triples, and the incidence
    lists (poly_start/poly_end over id_set)."""
    V = np.array([[x0 + i * h, y, z0 + k * h] for k in range(n) for i in range(n)])
    tris = []
    for k in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = k * n + i, k * n + i + 1, (k + 1) * n + i, (k + 1) * n + i + 1
            tris += [(a, c, b), (b, c, d)]  # (Q - P) x (R - P) points along +y
    tris = np.array(tris, np.uint32)
    inc = [[] for _ in range(len(V))]
    for t, (a, b, c) in enumerate(tris):
        for v in (a, b, c):
            inc[v].append(t)
    id_set = np.array([t for l in inc for t in l], np.uint32)
    pend = np.cumsum([len(l) for l in inc]).astype(np.uint32)
    pstart = (pend - np.array([len(l) for l in inc])).astype(np.uint32)
    return V, tris, id_set, pstart, pend


def _shape_case(orc, particles):
    r = float(scenes.as_f32_exact(3e-3))
    V, tris, id_set, pstart, pend = _tri_mesh()
    n = len(particles)
    pos = np.c_[np.array(particles), np.full(n, r)]
    sc = scenes.Scene(name="s", pos=pos, dt=1e-5, nsteps=1, world_origin=(0.0, -0.05, 0.0),
                      cell_size=2 * r, grid_size=(64, 64, 64))
    p = setb_params(sc, n, nsp=len(V))
    B = orc.SetB(p)
    data = B.mergedata(pos, V)
    h, idx = B.hash_and_index(data)
    cs, ce, srt = B.reorder(h, idx, p.ncell)
    polys = B.update_polygons_information(tris, V)
    mi = abi.BMassInfo(abi._d3((0, 0, 0)), abi._d3((0, 0, 0)))
    k = (5e5, 3.0, 4e5, 2.5, 0.3)
    vel = as4(np.zeros((n, 3)), 0.0)
    om = as4(np.zeros((n, 3)), 0.0)
    f = np.full((n, 3), 9.0)  # cleared by the call
    m = np.full((n, 3), 9.0)
    is_line, bf = B.shape(mi, polys, id_set, pstart, pend, k, np.ascontiguousarray(pos), vel, om,
                          f, m, srt, cs, ce, n)
    return dict(f=f, m=m, is_line=is_line, bf=bf, r=r, k=k, V=V)


def test_shape_face_edge_and_body_force(orc):
    r = float(scenes.as_f32_exact(3e-3))
    d = 2e-5
    cases = [[0.061, r - d, 0.063],           # above the mesh interior: face contact
             [0.05 - r + d, 0.0, 0.063],      # beside the x = 0.05 edge, in the plane: edge
             [0.061, 0.03, 0.063]]            # far above: nothing
    q = _shape_case(orc, cases)
    kn = q["k"][0]
    assert np.allclose(q["f"][0], [0, kn * d ** 1.5, 0], rtol=1e-9)
    assert np.allclose(q["f"][1], [-kn * d ** 1.5, 0, 0], rtol=1e-6)
    assert (q["f"][2] == 0).all() and (q["m"] == 0).all()
    assert list(q["is_line"]) == [True, True, False]
    assert np.allclose(q["bf"], q["f"].sum(0), rtol=1e-12, atol=1e-18)


def test_polygon_updates_closed_form(orc):
    A = np.array([0.0, -1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0])  # 90 deg about z
    lv = np.array([[1.0, 2.0, 3.0], [-4.0, 5.0, 0.5]])
    out = orc.SetB.update_polygons_position(np.array([10.0, 20.0, 30.0]), A, lv)
    assert np.array_equal(out, [[8.0, 21.0, 33.0], [5.0, 16.0, 30.5]])
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0], [0.0, 3.0, 0]])
    polys = orc.SetB.update_polygons_information(np.array([[0, 1, 2]]), pos)
    po = polys[0]
    assert (po.V.x, po.V.y, po.V.z) == (2.0, 0.0, 0.0)
    assert (po.W.x, po.W.y, po.W.z) == (0.0, 3.0, 0.0)
    assert (po.N.x, po.N.y, po.N.z) == (0.0, 0.0, 6.0)
    data = orc.SetB.mergedata(np.array([[1.0, 2, 3, 4]]), np.array([[5.0, 6, 7]]))
    assert np.array_equal(data, [[1, 2, 3, 4], [5, 6, 7, -1]])
