"""Drum walls and triangle meshes on the fused FP32c path against the FP64
oracle (the same scene the FP64 GPU path is tested on: a spinning drum with a
fin turning with it, particles against the shell, the caps and both faces of
the fin).  Bars: per-particle force m a of a step within FP32C_TOL (1e-5) of
the oracle, relative to the particle's summed contact-force magnitudes plus
its weight; the fin's reaction force and moment within 1e-5 of the summed
magnitudes of its contact forces (times the largest lever for the moment); runs alone vs in a mix, graph vs eager and through a
checkpoint bit-identical."""
import numpy as np
import pytest

from metakonga_amd import abi, scenes
from metakonga_amd.batch import Batch
from test_gpu_drum import _drum_scene
from xdem_testutil import FP32C_TOL, acc_error, world_step_magnitudes

pytestmark = pytest.mark.gpu


def _batch_for(scs):
    return Batch(sum(s.np for s in scs), len(scs), sum(int(np.prod(s.grid_size)) for s in scs))


def _acc_err(st, w, fm):
    err = acc_error(st["acc"], w, fm)
    print(f"FP32c vs FP64 oracle: force err {err}")
    return err["term"]


def test_first_step_vs_oracle(lib, orc):
    sc, vel = _drum_scene()
    b = _batch_for([sc])
    sid = b.add(sc, vel)
    b.graphs(False)
    b.step(1)
    st = b.state(sid)
    w = orc.World(sc)
    w.vel[:] = vel
    fm = world_step_magnitudes(orc, w)
    assert np.abs(st["pos"] - w.pos).max() < 1e-9
    assert _acc_err(st, w, fm) < FP32C_TOL
    b.close()


@pytest.mark.parametrize("start_steps", [15, 40])
def test_later_step_vs_oracle(lib, orc, start_steps):
    """From the fused state after some steps, one more step on both paths:
    the fin's pose at that step, the spinning walls and the particle contacts."""
    sc, vel = _drum_scene()
    b = _batch_for([sc])
    sid = b.add(sc, vel)
    b.step(start_steps)
    st = b.state(sid)
    w = orc.World(sc)
    for k in ("pos", "vel", "omega", "acc", "alpha"):
        getattr(w, k)[:] = st[k]
    w.w.step = start_steps
    for m in w.motions:
        m.n_act, m.active = start_steps, 1
    b.step(1)
    fm = world_step_magnitudes(orc, w)
    st2 = b.state(sid)
    assert np.abs(st2["pos"] - w.pos).max() < 1e-9
    assert _acc_err(st2, w, fm) < FP32C_TOL
    b.close()


@pytest.mark.parametrize("start_steps", [2, 10, 27])
def test_motion_window_vs_oracle(lib, orc, start_steps):
    """A fin that turns only from t = 5e-5 s to 2.5e-4 s (steps 5..25 at dt
    1e-5): one step from the fused state before, during and after its motion
    against the oracle (pose advanced by the active steps so far, velocity zero
    outside them)."""
    start, end = 5e-5, 2.5e-4
    sc, vel = _drum_scene(start=start, end=end)
    b = _batch_for([sc])
    sid = b.add(sc, vel)
    b.step(start_steps)
    st = b.state(sid)
    w = orc.World(sc)
    for k in ("pos", "vel", "omega", "acc", "alpha"):
        getattr(w, k)[:] = st[k]
    w.w.step = start_steps
    act = [k * sc.dt >= start and k * sc.dt <= end for k in range(1, start_steps + 1)]
    for m in w.motions:
        # active = 1 makes the oracle re-pose the fin at the next step even when
        # the motion is off then (as after the transition the fused run went through)
        m.n_act, m.active = sum(act), 1
    b.step(1)
    fm = world_step_magnitudes(orc, w)
    st2 = b.state(sid)
    assert np.abs(st2["pos"] - w.pos).max() < 1e-9
    assert _acc_err(st2, w, fm) < FP32C_TOL
    b.close()


def _run(scs, vels, steps, block, graphs=True):
    b = _batch_for(scs)
    b.graphs(graphs)
    ids = [b.add(s, v) for s, v in zip(scs, vels)]
    done = 0
    while done < steps:
        b.step(min(block, steps - done))
        done += block
    out = [b.state(i) for i in ids]
    b.close()
    return out


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_alone_vs_mix_and_graphs(lib):
    sc, vel = _drum_scene()
    box = scenes.box_scene(1500, 3e-3, 91, 1e-5, 100)
    alone = _run([sc], [vel], 30, 10)[0]
    _same(alone, _run([box, sc], [None, vel], 30, 10)[1])
    _same(alone, _run([sc], [vel], 30, 10, graphs=False)[0])


@pytest.mark.parametrize("window", [(0.0, 0.0), (5e-5, 2.5e-4)])
def test_checkpoint_resume(lib, tmp_path, window):
    """Bit-identical through a checkpoint at step 15, also with the fin's motion
    window open then (steps 5..25)."""
    sc, vel = _drum_scene(start=window[0], end=window[1])
    ref = _run([sc], [vel], 30, 15)[0]
    b = _batch_for([sc])
    b.add(sc, vel)
    b.step(15)
    b.save_checkpoint(tmp_path / "g.ckpt")
    b.close()
    b2 = _batch_for([sc])
    (sid,) = b2.load_checkpoint(tmp_path / "g.ckpt")
    b2.step(15)
    _same(ref, b2.state(sid))
    b2.close()


def _fin_scale(sc, w):
    """Summed magnitudes of the fin particles' contact forces (those particles
    touch nothing else: placed against the fin, clear of shell and lattice)."""
    g = np.asarray(sc.gravity)
    fc = w.mass * np.linalg.norm(w.acc - g[None, :], axis=1)   # a = (m g + F) / m
    fin = (np.abs(w.pos[:, 0]) < 1.2 * 2e-3) & (w.pos[:, 1] < -0.026)
    assert fin.sum() >= 5 and (fc[fin] > 0).all()
    return fc[fin].sum()


@pytest.mark.parametrize("start_steps", [0, 25])
def test_reactions_vs_oracle(lib, orc, start_steps):
    """The fin's reaction (minus its summed triangle contact forces, and their
    moment about the fin origin) against the oracle's dpmi force and moment of
    the same step; the drum's reaction against the oracle's drum reactions."""
    sc, vel = _drum_scene()
    b = _batch_for([sc])
    sid = b.add(sc, vel)
    b.enable_reactions(True)
    w = orc.World(sc)
    w.vel[:] = vel
    if start_steps:
        b.step(start_steps)
        st = b.state(sid)
        for k in ("pos", "vel", "omega", "acc", "alpha"):
            getattr(w, k)[:] = st[k]
        w.w.step = start_steps
        for m in w.motions:
            m.n_act, m.active = start_steps, 1
    vel_in, om_in = w.vel.copy(), w.omega.copy()
    b.step(1)
    w.step(1)
    # the drum: the oracle's per-particle drum reactions at contact time
    # (positions after the step's position update, the step's input velocities)
    mf, mm = orc.drum(sc.params(), sc.tcm, sc.drums[0], w.pos, vel_in, om_in,
                      np.zeros((sc.np, 3)), np.zeros((sc.np, 3)), w.mass, sc.contact_pp(),
                      reaction=True)
    df, dm = b.drum_reactions(sid)
    dscale = np.linalg.norm(mf, axis=1).sum()
    assert df.shape == (1, 3) and dscale > 0
    assert np.abs(df[0] - mf.sum(0)).max() <= FP32C_TOL * dscale, (df[0], mf.sum(0))
    assert np.abs(dm[0] - mm.sum(0)).max() <= FP32C_TOL * dscale * 0.05, (dm[0], mm.sum(0))
    assert abs(mm.sum(0)[2]) > 1e-3 * dscale * 0.04     # a drive torque about the axis
    f, m = b.object_reactions(sid)
    assert f.shape == (1, 3) and m.shape == (1, 3)
    d = w.dpmi[0]
    of = np.array([d.force.x, d.force.y, d.force.z])
    om = np.array([d.moment.x, d.moment.y, d.moment.z])
    scale = _fin_scale(sc, w)
    assert np.abs(of).max() > 1e-3 * scale            # a net reaction to compare
    assert np.abs(f[0] - of).max() <= FP32C_TOL * scale, (f[0], of, scale)
    assert np.abs(m[0] - om).max() <= FP32C_TOL * scale * 0.05, (m[0], om)
    b.close()


def test_reactions_deterministic_and_errors(lib):
    sc, vel = _drum_scene()
    box = scenes.box_scene(1500, 3e-3, 91, 1e-5, 100)
    outs = []
    for scs, vels, graphs in (([sc], [vel], True), ([box, sc], [None, vel], True),
                              ([sc], [vel], False)):
        b = _batch_for(scs)
        b.graphs(graphs)
        ids = [b.add(s, v) for s, v in zip(scs, vels)]
        b.enable_reactions(True)
        with pytest.raises(RuntimeError):      # nothing computed yet
            b.object_reactions(ids[-1])
        b.step(10)
        b.step(10)
        outs.append(b.object_reactions(ids[-1]))
        if len(ids) == 2:
            f0, m0 = b.object_reactions(ids[0])   # a scene without objects or drums
            assert f0.shape == (0, 3) and b.drum_reactions(ids[0])[0].shape == (0, 3)
        outs.append(b.drum_reactions(ids[-1]))
        b.enable_reactions(False)
        with pytest.raises(RuntimeError):
            b.object_reactions(ids[-1])
        b.close()
    for kind in (outs[0::2], outs[1::2]):         # objects, drums
        for f, m in kind[1:]:
            assert np.array_equal(f, kind[0][0]) and np.array_equal(m, kind[0][1])


def test_refusals(lib):
    sc, vel = _drum_scene()
    b = _batch_for([sc])
    sid = b.add(sc, vel)
    b.step(1)
    dr = sc.drums[0]
    wall = abi.ContactPairDesc(0.8, 0.3, 0.0, 0.0, 0.8, abi.Material(2500, 1e7, 0.3, 0))
    import ctypes as C
    with pytest.raises(RuntimeError):       # the scene has stepped
        abi.check(lib, lib.xdem_batch_add_drum(b.h, sid, C.byref(dr), C.byref(wall)))
    b.close()
    small = scenes.Scene(name="s", pos=sc.pos, dt=sc.dt, nsteps=1, world_origin=sc.world_origin,
                         cell_size=4.1e-3, grid_size=sc.grid_size, tri=sc.tri,
                         objects=sc.objects)
    b = _batch_for([small])
    with pytest.raises(RuntimeError):       # 4.1 mm < r_max + circumradius = 4.12 mm
        b.add(small)
    b.close()
    # tangential history: with drums yes (drum springs), with meshes no
    hist = scenes.Scene(name="h", pos=sc.pos, dt=sc.dt, nsteps=1, world_origin=sc.world_origin,
                        cell_size=0.0, grid_size=sc.grid_size, drums=sc.drums,
                        tangential_history=True)
    b = _batch_for([hist])
    hid = b.add(hist)
    b.enable_reactions(True)
    b.step(1)
    with pytest.raises(RuntimeError):       # drum reactions are evaluated without springs
        b.drum_reactions(hid)
    b.close()
    mesh = scenes.Scene(name="hm", pos=sc.pos, dt=sc.dt, nsteps=1, world_origin=sc.world_origin,
                        cell_size=0.0, grid_size=sc.grid_size, tri=sc.tri, objects=sc.objects,
                        tangential_history=True)
    b = _batch_for([mesh])
    with pytest.raises(RuntimeError):
        b.add(mesh)
    b.close()


def test_c2_full_size_properties(lib):
    """The drum workload at full size (4e6 spheres, a 20,000-triangle lifter
    plate turning with the drum): two runs
    are bit-identical, stay finite and keep every particle inside the drum."""
    sc = scenes.drum()
    outs = []
    for _ in range(2):
        b = Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
        sid = b.add(sc)
        b.step(20)
        b.step(20)
        outs.append(b.state(sid))
        b.close()
    _same(outs[0], outs[1])
    st = outs[0]
    assert np.isfinite(st["pos"]).all() and np.isfinite(st["vel"]).all()
    rho = np.hypot(st["pos"][:, 0], st["pos"][:, 1])
    assert rho.max() < 0.5 and np.abs(st["pos"][:, 2]).max() < 0.8
    assert (np.abs(st["vel"]).max(1) > 0).all()   # gravity acts on every sphere
