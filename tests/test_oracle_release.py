"""Real-time particle release in the oracle (dem_simulation::oneStepAnalysis,
source/simulation/dem_simulation.cpp:449-450: at step k with (k - 1) % per_np
== 0 the next group joins the simulated prefix) against the schedule and the
free-fall closed form of the GPU-form Verlet update."""
import numpy as np
import pytest

from metakonga_amd import scenes

G = np.array([0.0, -9.80665, 0.0])


def _column(n, r=2e-3, gap=0.05):
    """n particles far apart in a vertical column (no contacts, no walls)."""
    y = 0.5 + gap * np.arange(n)
    return np.c_[np.zeros(n), y, np.zeros(n), np.full(n, r)]


def test_release_schedule_and_free_fall(orc):
    dt, every, group, n = 1e-4, 5, 4, 10
    pos0 = _column(n)
    sc = scenes.Scene(name="rel", pos=pos0.copy(), dt=dt, nsteps=20,
                      world_origin=(-1.0, -1.0, -1.0), cell_size=0.01, grid_size=(256, 256, 256),
                      release=(every, group))
    w = orc.World(sc)
    assert w.w.np == 0
    for k in range(1, 17):
        w.step(1)
        assert w.w.np == min(n, group * ((k - 1) // every + 1)), k
        for i in range(n):
            first = (i // group) * every + 1          # the step the particle joins
            m = k - first + 1                         # steps it has been simulated
            if m <= 0:
                assert np.array_equal(w.pos[i], pos0[i]), (k, i)   # untouched until released
                continue
            # GPU-form Verlet from rest with a = 0 at the first step:
            # y = y0 + g (T^2 - T dt) / 2, T = m dt
            T = m * dt
            assert abs(w.pos[i, 1] - (pos0[i, 1] + 0.5 * G[1] * (T * T - T * dt))) < 1e-14, (k, i)
            assert abs(w.vel[i, 1] - (G[1] * T - 0.5 * G[1] * dt)) < 1e-12, (k, i)


def test_release_group_list(orc):
    """A list of group sizes (particleManager::NextCreatingPerGroup,
    source/particleManager.cpp:132-139): the g-th release adds groups[g]
    particles and none after the list ends, so a particle beyond the list's
    total is never simulated."""
    dt, every, groups, n = 1e-4, 3, [3, 4, 2], 10
    pos0 = _column(n)
    sc = scenes.Scene(name="rel", pos=pos0.copy(), dt=dt, nsteps=20,
                      world_origin=(-1.0, -1.0, -1.0), cell_size=0.01, grid_size=(256, 256, 256),
                      release=(every, groups))
    w = orc.World(sc)
    first = np.repeat(np.arange(len(groups)) * every + 1, groups)   # the step each one joins
    for k in range(1, 15):
        w.step(1)
        assert w.w.np == sum(groups[:(k - 1) // every + 1]), k
        for i in range(n):
            m = k - first[i] + 1 if i < len(first) else 0
            if m <= 0:
                assert np.array_equal(w.pos[i], pos0[i]), (k, i)
                continue
            T = m * dt
            assert abs(w.pos[i, 1] - (pos0[i, 1] + 0.5 * G[1] * (T * T - T * dt))) < 1e-14, (k, i)
    assert w.w.np == 9 and w.w.release_next == 3


def test_release_period():
    """dem_simulation.cpp:213,216: (unsigned)((t + 1e-9) / dt) for groups,
    (unsigned)((1.0 / per_second) / dt) one at a time.  0.01 / 1e-5 rounds to
    999.999..., so the group form's 1e-9 guard gives 1000 while the
    per-second form of the same period truncates to 999."""
    assert scenes.release_period(1e-5, time_per_group=0.01) == 1000
    assert scenes.release_period(1e-5, per_second=100) == 999
    assert scenes.release_period(1e-3, per_second=3) == 333
    assert scenes.release_period(1e-4, time_per_group=0.009) == 90
    assert scenes.release_period(1e-3, time_per_group=1e-4) == 0      # no release
    with pytest.raises(ValueError):
        scenes.release_period(1e-3)
    with pytest.raises(ValueError):
        scenes.release_period(1e-3, per_second=0)


def _xdm(shape, making):
    gen = {"plane": "DIMENSION 0.1 3 0.06\n",
           "circle": "DIAMETER 0.05\nNUM_HEIGHT 2\n"}[shape]
    return f"""MODEL_NAME rel
GRAVITY 0 -9.80665 0
UNIT 0
SINGLE 0

DEM_MODEL_DATA
INTEGRATOR 0
SORT 0

PARTICLES_DATA
CREATE_SHAPE {shape}
NAME particles
MATERIAL_TYPE 0
{gen}STARTPOINT 0 0.1 0
DIRECTION 0 1 0
SPACE 0
RADIUS 0.004 0.004
MAKING {making} none
MATERIAL 10000000 2500 0.3 3846153.8
END_DATA
"""


def test_xdm_release():
    """`MAKING realtime np per_np time_per_group one_by_one file` of the last
    plane/circle generator (particleManager.cpp:392-397, :588-594): groups
    release the generators' layers; one-by-one releases single particles at
    per_np a second; a plane generator's one-by-one rate divides by zero in
    the reference and is refused."""
    from metakonga_amd import model_io as mio
    m = mio.parse_xdm(_xdm("plane", "1 0 0 0.01 0"))
    sc = m.to_scene(dt=1e-5, nsteps=10)
    _, groups = mio.plane_particles(0.1, 3, 0.06, (0, 0.1, 0), 0.004, 0.004)
    assert sc.release[0] == 1000 and list(sc.release[1]) == groups and sum(groups) == sc.np
    assert mio.parse_xdm(_xdm("plane", "0 0 0 0.01 0")).to_scene(1e-5, 10).release is None
    with pytest.raises(ValueError):
        mio.parse_xdm(_xdm("plane", "1 0 50 0 1")).to_scene(1e-5, 10)
    sc = mio.parse_xdm(_xdm("circle", "1 0 50 0 1")).to_scene(1e-4, 10)
    assert sc.release == (200, 1)
    sc = mio.parse_xdm(_xdm("circle", "1 0 0 0.002 0")).to_scene(1e-4, 10)
    _, groups = mio.circle_particles(0.05, 2, (0, 0.1, 0), 0.004, 0.004)
    assert sc.release[0] == 20 and list(sc.release[1]) == groups and len(groups) == 2
