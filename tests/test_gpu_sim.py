"""Whole-step parity: dem_simulation::oneStepAnalysis on the GPU (xdem_sim, plain
FP64) against the oracle's step, on a contact-rich box scene."""
import numpy as np
import pytest

from metakonga_amd import scenes
from metakonga_amd.xdem import Simulation
from xdem_testutil import dense_state

pytestmark = pytest.mark.gpu


def _scene():
    pos, vel, om = dense_state(n_side=10, pitch_factor=1.99, jitter=0.03)
    r = pos[0, 3]
    lo = pos[:, :3].min(0) - r
    hi = pos[:, :3].max(0) + r
    sc = scenes.Scene(name="dense-box", pos=pos, dt=1e-5, nsteps=30,
                      world_origin=tuple(lo - 2 * r), cell_size=2 * r, grid_size=(32, 32, 32),
                      planes=scenes.open_box_planes(lo + 0.2 * r, hi + [0, 0.05, 0]))
    return sc, vel, om


def test_steps_match_oracle(lib, orc):
    sc, vel, om = _scene()
    sim = Simulation(sc, lib)
    m, I = sim.mass_inertia()
    w = orc.World(sc, mass=m, iner=I)
    w.vel[:] = vel
    w.omega[:] = om
    sim.set_state(vel=vel, omega=om)
    for k in range(3):
        sim.step(10)
        w.step(10)
        st = sim.state()
        scale = np.abs(w.pos[:, :3]).max()
        dpos = np.abs(st["pos"] - w.pos).max() / scale
        dvel = np.abs(st["vel"] - w.vel).max() / (np.abs(w.vel).max() + 1e-30)
        assert dpos < 1e-13, (k, dpos)
        assert dvel < 1e-9, (k, dvel)
    sim.close()


def test_run_twice_bitexact(lib):
    """Determinism: no atomics on the contact path, fixed reduction order."""
    sc, vel, om = _scene()
    outs = []
    for _ in range(2):
        sim = Simulation(sc, lib)
        sim.set_state(vel=vel, omega=om)
        sim.step(25)
        outs.append(sim.state())
        sim.close()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def _release_scene(release=(40, 150)):
    """Three groups of 150 spheres released every 40 steps into an open box,
    stacked: the first on the floor, each in 3 slightly overlapping layers of
    50, the later groups above the earlier ones (dem_simulation.cpp:449-450).
    Every particle gets a random velocity and spin, which a group keeps until
    its release: a pair that starts from rest with one contact has a relative
    velocity along its line of centres, and the reference's tangential
    direction e/|e| is then rounding noise that the Coulomb branch scales to
    mu*|fsn| (mphysics_cuda_impl.cuh:379-401), so the result would hang on
    how each compiler rounds rather than on the law.  dt = 5e-6 keeps the
    explicit tangential dashpot (coefficient ks dt + vs on |e|) stable for the
    spin of a 2 mm sphere: at 2e-5 sliding contacts chatter and any rounding
    difference grows by e every few steps."""
    r = 2e-3
    g = (np.arange(10) + 0.5) * 2.2 * r
    X, Z = np.meshgrid(g, g, indexing="ij")
    L = np.c_[X.ravel(), np.zeros(X.size), Z.ravel()][:50]
    rng = np.random.default_rng(3)
    pos = []
    for k in range(3):
        base = r - 1e-5 + k * (3 * 2 * r + 1e-3)
        for h in range(3):
            P = L + [0.0, base + h * (2 * r - 2e-6), 0.0]
            P[:, [0, 2]] += rng.uniform(-0.05 * r, 0.05 * r, (len(L), 2))
            pos.append(P)
    pos = np.concatenate(pos)
    pos = np.c_[pos, np.full(len(pos), r)]
    planes = scenes.open_box_planes((-0.002, 0.0, -0.002), (0.05, 0.05, 0.05))
    sc = scenes.Scene(name="release", pos=np.ascontiguousarray(pos), dt=5e-6, nsteps=120,
                      world_origin=(-0.05, -0.05, -0.05), cell_size=2 * r,
                      grid_size=(64, 64, 64), planes=planes, release=release)
    u = rng.uniform(-1.0, 1.0, (len(pos), 6))
    return sc, scenes.as_f32_exact(0.05 * u[:, :3]), scenes.as_f32_exact(5.0 * u[:, 3:])


# (release, [(steps, particles active after them)]): a constant group, and a
# list of unequal groups (particleManager::np_group, particleManager.cpp:132-139)
RELEASES = [((40, 150), [(1, 150), (39, 150), (1, 300), (60, 450), (19, 450)]),
            ((40, [100, 150, 200]), [(1, 100), (39, 100), (1, 250), (39, 250), (1, 450),
                                     (39, 450)])]


@pytest.mark.parametrize("release,schedule", RELEASES, ids=["group", "group-list"])
def test_release_matches_oracle(lib, orc, release, schedule):
    sc, vel, om = _release_scene(release)
    sim = Simulation(sc, lib)
    m, I = sim.mass_inertia()
    w = orc.World(sc, mass=m, iner=I)
    w.vel[:] = vel
    w.omega[:] = om
    sim.set_state(vel=vel, omega=om)
    assert sim.active == 0
    scale = np.abs(w.pos[:, :3]).max()
    for k, (n, active) in enumerate(schedule):
        sim.step(n)
        w.step(n)
        assert sim.active == active == w.w.np, k
        st = sim.state()
        dpos = np.abs(st["pos"] - w.pos).max() / scale
        dvel = np.abs(st["vel"] - w.vel).max() / np.abs(w.vel).max()
        print(f"[release] after {sim.steps} steps: active {active} dpos {dpos:.3e} dvel {dvel:.3e}",
              flush=True)
        assert dpos < 1e-13, (k, dpos)
        assert dvel < 1e-9, (k, dvel)
        # particles not yet released have not moved
        assert np.array_equal(st["pos"][active:], sc.pos[active:]), k
        assert np.array_equal(st["vel"][active:], vel[active:]), k
    assert sim.steps == 120
    sim.close()


@pytest.mark.parametrize("release,active60", [((40, 150), 300), ((40, [100, 150, 200]), 250)],
                         ids=["group", "group-list"])
def test_release_checkpoint_resumes_bitexact(lib, tmp_path, release, active60):
    sc, vel, om = _release_scene(release)
    ref = Simulation(sc, lib)
    ref.set_state(vel=vel, omega=om)
    ref.step(60)
    path = tmp_path / "rel.ckpt"
    ref.save_checkpoint(path)
    ref.step(50)
    sim = Simulation(sc, lib)
    sim.load_checkpoint(path)
    assert sim.active == active60
    sim.step(50)
    a, b = ref.state(), sim.state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert sim.active == ref.active == 450
    ref.close()
    sim.close()
