"""Pins of the oracle's cohesion term (CPU only).

cohesionForce (include/mphysics_cuda_impl.cuh:259-279, called from HMCModel
:357 and DHSModel :386) and its CPU twin contact::cohesionForce
(source/collision/contact.cpp:6-18) return coh * pi * rc^2 with
rc^3 = 3 R* |Fn| / (4 E*): rc is the Hertz contact radius for the normal load
|Fn|.  For a Hertz contact (|Fn| = kn delta^1.5 with kn = 4/3 E* sqrt(R*))
the contact radius is sqrt(R* delta), so the cohesive force is the cohesive
stress times the contact area, coh * pi * R* * delta.  That identity, a hand
known answer and the sign in a two-sphere contact pin the restatement; a
dropped factor (3/4), a wrong exponent or E instead of E* fails them.
"""
import math

import numpy as np
import pytest

from metakonga_amd import abi, scenes
from test_oracle_forces import CP, MAT, _pair_state, estar, prm


def eeq_sum(Ei, Ej, nui, nuj):
    return (1 - nui * nui) / Ei + (1 - nuj * nuj) / Ej


def test_known_answer(orc):
    """r = 5 mm spheres of E = 1e7, nu = 0.3 under 1 N: Hertz contact radius
    a = (3 F R* / (4 E*))^(1/3) = 6.988075e-4 m, area 1.534140e-6 m^2."""
    cf = orc.cohesion_force(5e-3, 5e-3, 1e7, 1e7, 0.3, 0.3, 1e4, -1.0)
    assert cf == pytest.approx(0.0153413995, rel=1e-9)
    # the CPU form with coh_r = R*, coh_e = sum (1 - nu^2) / E
    ch = orc.host_cohesion(1e4, 2.5e-3, eeq_sum(1e7, 1e7, 0.3, 0.3), -1.0)
    assert ch == pytest.approx(cf, rel=1e-15)


@pytest.mark.parametrize("ri,rj,Ei,Ej,nui,nuj", [
    (5e-3, 5e-3, 1e7, 1e7, 0.3, 0.3),
    (2e-3, 4e-3, 1e7, 1e7, 0.3, 0.3),
    (3e-3, 1.1e-3, 2e8, 5e6, 0.25, 0.41),
])
@pytest.mark.parametrize("delta", [1e-7, 3.3e-5, 4e-4])
def test_hertz_contact_area(orc, ri, rj, Ei, Ej, nui, nuj, delta):
    """With the getConstant stiffness the cohesion is coh pi R* delta."""
    coh = 2.5e3
    R = ri * rj / (ri + rj)
    Es = 1.0 / eeq_sum(Ei, Ej, nui, nuj)
    kn = 4.0 / 3.0 * Es * math.sqrt(R)
    cp = abi.ContactProperty(Ei, Ej, nui, nuj, 0, 0, 0.8, 0.3, 0.0, coh, 0.8)
    c = orc.force_constant(1, ri, rj, 1.0, 1.0, cp)
    assert c.kn == pytest.approx(kn, rel=1e-14)
    fsn = -c.kn * delta ** 1.5
    cf = orc.cohesion_force(ri, rj, Ei, Ej, nui, nuj, coh, fsn)
    assert cf == pytest.approx(coh * math.pi * R * delta, rel=1e-13)
    ch = orc.host_cohesion(coh, R, eeq_sum(Ei, Ej, nui, nuj), fsn)
    assert ch == pytest.approx(cf, rel=1e-14)


def test_zero_cases(orc):
    assert orc.cohesion_force(5e-3, 5e-3, 1e7, 1e7, 0.3, 0.3, 0.0, -3.0) == 0.0
    assert orc.host_cohesion(0.0, 2.5e-3, 1.8e-7, -3.0) == 0.0
    assert orc.cohesion_force(5e-3, 5e-3, 1e7, 1e7, 0.3, 0.3, 1e4, 0.0) == 0.0
    # cohesion grows with the load as |Fn|^(2/3)
    a = orc.cohesion_force(5e-3, 5e-3, 1e7, 1e7, 0.3, 0.3, 1e4, -1.0)
    b = orc.cohesion_force(5e-3, 5e-3, 1e7, 1e7, 0.3, 0.3, 1e4, -8.0)
    assert b == pytest.approx(4.0 * a, rel=1e-14)


def _coh_p2p(orc, pos, vel, om, coh, semantics):
    n = pos.shape[0]
    p = prm(np_=n)
    h, idx = orc.hash_and_index(p, pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    m, I = scenes.mass_inertia(pos[:, 3], MAT.density)
    f = np.zeros((n, 3))
    t = np.zeros((n, 3))
    cp = scenes.contact_property(MAT, MAT, scenes.ContactSpec(cohesion=coh))
    orc.p2p(p, 1, pos, vel, om, f, t, m, srt, cs, ce, cp, semantics=semantics)
    return f, t


@pytest.mark.parametrize("semantics", [0, 1])
@pytest.mark.parametrize("d", [2e-6, 4.6335e-5])
def test_pair_at_rest(orc, semantics, d):
    """Two spheres at rest, overlap d: the normal force on sphere 0 is
    (-kn d^1.5 + coh pi R* d) along +x (towards sphere 1), on both the GPU
    kernel's and the CPU path's semantics; no tangential force, no torque."""
    r, coh = 5e-3, 3e3
    pos, vel, om = _pair_state(d)
    f, t = _coh_p2p(orc, pos, vel, om, coh, semantics)
    kn = 4.0 / 3.0 * estar() * math.sqrt(r / 2)
    want = -kn * d ** 1.5 + coh * math.pi * (r / 2) * d
    assert f[0, 0] == pytest.approx(want, rel=1e-12)
    assert f[1, 0] == pytest.approx(-want, rel=1e-12)
    assert np.all(f[:, 1:] == 0) and np.all(t == 0)
    # cohesion pulls: the net normal force is smaller than the Hertz repulsion
    f0, _ = _coh_p2p(orc, pos, vel, om, 0.0, semantics)
    assert f0[0, 0] == pytest.approx(-kn * d ** 1.5, rel=1e-12)
    assert f[0, 0] > f0[0, 0]


def test_walls_carry_no_cohesion(orc):
    """A wall contact has no cohesion term (the reference's wall kernels pass
    ri = rj = 0 and get 0/0; a recorded departure)."""
    r = 5e-3
    pl = scenes.plane4((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1))
    pos = np.array([[0.0, r - 3e-5, 0.0, r]])
    out = []
    for coh in (0.0, 5e3):
        f = np.zeros((1, 3))
        t = np.zeros((1, 3))
        m, I = scenes.mass_inertia(pos[:, 3], MAT.density)
        cp = scenes.contact_property(MAT, MAT, scenes.ContactSpec(cohesion=coh))
        orc.plane(prm(np_=1), 1, pl, pos, np.zeros((1, 3)), np.zeros((1, 3)), f, t, m, cp)
        out.append(f)
    assert out[0][0, 1] != 0
    assert np.array_equal(out[0], out[1])


def test_force_magnitudes(orc):
    """The magnitude accumulator adds |Fn + Ft| of every contact of a particle."""
    d = 3e-5
    pos, vel, om = _pair_state(d, vrel=(0, 0, 0.2))
    with orc.force_magnitudes(2) as fm:
        f, t = _coh_p2p(orc, pos, vel, om, 0.0, 0)
    assert fm[0] == pytest.approx(np.linalg.norm(f[0]), rel=1e-14)
    assert fm[1] == pytest.approx(np.linalg.norm(f[1]), rel=1e-14)
    # outside the block nothing is accumulated
    f2, _ = _coh_p2p(orc, pos, vel, om, 0.0, 0)
    assert np.array_equal(f, f2) and fm[0] == pytest.approx(np.linalg.norm(f[0]), rel=1e-14)
