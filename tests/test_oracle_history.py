"""Pins of the oracle's tangential-spring history (CPU).

The reference has no contact history (its DHS tangential force is memoryless,
mphysics_cuda_impl.cuh:379-401), so the history model is pinned by:
  * reduction to the reference law: with an empty history every contact's
    force equals the reference DHS force (|e| != 0 in these states);
  * keys: the per-particle key list equals the reference pair list in
    traversal order, followed by the touching planes in order (bit-exact);
  * a closed form: under a constant relative tangential velocity the spring
    grows linearly, xi_K = K dt e, Ft = ks xi + vs e, until the Coulomb cap
    mu |Fn|, where Ft saturates along e;
  * separation zeroes the spring (the key disappears and a new contact
    starts from zero).
"""
import numpy as np

from metakonga_amd import scenes
from xdem_testutil import dense_scene


def _world(orc, sc, vel, om):
    w = orc.World(sc)
    w.vel[:] = vel
    w.omega[:] = om
    w.detect()
    return w


def _plane_keys(orc, sc, pos):
    """Touching planes per particle (exact FP64 predicate), in order."""
    out = [[] for _ in range(sc.np)]
    for q, pl in enumerate(sc.planes):
        for i in range(sc.np):
            xw = np.array([pl.xw.x, pl.xw.y, pl.xw.z])
            dp = pos[i, :3] - xw
            wp = [dp @ [pl.u1.x, pl.u1.y, pl.u1.z], dp @ [pl.u2.x, pl.u2.y, pl.u2.z],
                  dp @ [pl.uw.x, pl.uw.y, pl.uw.z]]
            c, _ = orc.plane_detect(pl, pos[i, :3], wp, pos[i, 3])
            if c > 0:
                out[i].append(orc.ORC_WALL_KEY | q)
    return out


def test_empty_history_equals_reference_law(orc):
    sc, vel, om = dense_scene(n_side=7, seed=21)  # the outer layer touches the walls
    w = _world(orc, sc, vel, om)
    hin, hout = orc.History(sc.np), orc.History(sc.np)
    f, m = w.hist_contacts(hin, hout)
    p = sc.params()
    fr = np.zeros((sc.np, 3))
    mr = np.zeros((sc.np, 3))
    mass, _ = sc.mass_inertia()
    pairs = orc.p2p(p, 1, w.pos, w.vel, w.omega, fr, mr, mass, w.sorted, w.cstart, w.cend,
                    sc.contact_pp(), want_pairs=True)
    for pl in sc.planes:
        orc.plane(p, 1, pl, w.pos, w.vel, w.omega, fr, mr, mass, sc.contact_pp())
    assert len(pairs) > 3 * sc.np
    scale = np.abs(fr).max() + 1e-30
    assert np.abs(f - fr).max() / scale < 1e-12
    assert np.abs(m - mr).max() / (np.abs(mr).max() + 1e-30) < 1e-12
    # keys: particle partners in traversal order, then planes in order
    pk = _plane_keys(orc, sc, w.pos)
    assert sum(len(k) for k in pk) > 0
    for i in range(sc.np):
        want = list(pairs[pairs[:, 0] == i, 1]) + pk[i]
        keys, xi = hout.entries(i)
        assert list(keys) == want, i
    assert hout.overflow == 0


def _two_sphere_scene(dv_t, overlap=2e-5, mu=0.3):
    r = float(scenes.as_f32_exact(5e-3))
    pos = np.array([[0.0, 0.0, 0.0, r], [2 * r - overlap, 0.0, 0.0, r]])
    sc = scenes.Scene(name="pair", pos=pos, dt=1e-5, nsteps=1, world_origin=(-0.08, -0.08, -0.08),
                      cell_size=2 * r, grid_size=(16, 16, 16), gravity=(0.0, 0.0, 0.0),
                      rolling=False,
                      contact=scenes.ContactSpec(friction=mu))
    vel = np.array([[0.0, 0.0, 0.0], [0.0, dv_t, 0.0]])
    return sc, vel


def test_constant_tangential_velocity_closed_form(orc):
    dv_t = 1e-5
    sc, vel = _two_sphere_scene(dv_t)
    w = _world(orc, sc, vel, np.zeros((2, 3)))
    mass, _ = sc.mass_inertia()
    c = orc.force_constant(1, sc.pos[0, 3], sc.pos[1, 3], mass[0], mass[1], sc.contact_pp())
    delta = 2 * sc.pos[0, 3] - (sc.pos[1, 0] - sc.pos[0, 0])
    fn = c.kn * delta ** 1.5  # no normal velocity: |Fn| = kn delta^1.5
    mu = sc.contact_pp().fric
    e = np.array([0.0, dv_t, 0.0])  # relative tangential velocity of 1 seen from 0
    h = [orc.History(2), orc.History(2)]
    saturated_at = None
    for K in range(1, 4000):
        f, _ = w.hist_contacts(h[0], h[1])
        keys, xi = h[1].entries(0)
        assert list(keys) == [1]
        ft = f[0] - f[0] @ np.array([1.0, 0, 0]) * np.array([1.0, 0, 0])
        lin = (c.ks * K * sc.dt + c.vs) * dv_t
        if lin < mu * fn * (1 - 1e-9):
            assert np.allclose(xi[0], K * sc.dt * e, rtol=1e-12, atol=0)
            assert np.allclose(ft, lin * e / dv_t, rtol=1e-12, atol=1e-18)
        elif lin > mu * fn * (1 + 1e-9):
            saturated_at = saturated_at or K
            assert abs(np.linalg.norm(ft) - mu * fn) < 1e-12 * mu * fn
            assert ft[1] > 0 and abs(ft[2]) < 1e-15
            # the spring is held at the cap: ks xi + vs e = mu |Fn| e_hat
            assert abs(c.ks * xi[0][1] + c.vs * dv_t - mu * fn) < 1e-12 * mu * fn
        h = h[::-1]
        if saturated_at and K > saturated_at + 5:
            break
    assert saturated_at is not None and saturated_at > 10


def test_separation_zeroes_the_spring(orc):
    sc, vel = _two_sphere_scene(1e-3)
    w = _world(orc, sc, vel, np.zeros((2, 3)))
    h = [orc.History(2), orc.History(2)]
    for _ in range(5):
        w.hist_contacts(h[0], h[1])
        h = h[::-1]
    assert h[0].count[0] == 1 and np.abs(h[0].xi[0, 0]).max() > 0
    w.pos[1, 0] += 1e-4  # apart
    w.detect()
    w.hist_contacts(h[0], h[1])
    assert h[1].count[0] == 0 and h[1].count[1] == 0
    h = h[::-1]
    w.pos[1, 0] -= 1e-4  # touching again: the spring restarts from zero
    w.detect()
    w.hist_contacts(h[0], h[1])
    keys, xi = h[1].entries(0)
    assert list(keys) == [1]
    assert np.allclose(xi[0], sc.dt * np.array([0.0, 1e-3, 0.0]), rtol=1e-12)


def test_history_slots_overflow_is_counted(orc):
    """A particle with more contacts than slots keeps the first ones (in
    evaluation order) and counts the rest."""
    sc, vel, om = dense_scene(n_side=5, seed=4, pitch=1.7, jitter=0.02)
    w = _world(orc, sc, vel, om)
    small_in, small_out = orc.History(sc.np, slots=3), orc.History(sc.np, slots=3)
    big_in, big_out = orc.History(sc.np), orc.History(sc.np)
    w.hist_contacts(small_in, small_out)
    w.hist_contacts(big_in, big_out)
    total = int(big_out.count.sum())
    kept = int(small_out.count.sum())
    assert small_out.overflow == total - kept > 0
    for i in range(sc.np):
        c = int(small_out.count[i])
        assert np.array_equal(small_out.key[i, :c], big_out.key[i, :c])
