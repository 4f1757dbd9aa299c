"""State files on the GPU paths: checkpoints resume bit-identically (FP64
simulation with a spinning drum, a turning baffle and a rising floor; batched
FP32c scenes with and without tangential history), and the reference's part
and .bfr files carry the current state (a .bfr resume restarts from zero
accelerations and F = m g, as dem_simulation::initialize + setStartingData,
source/simulation/dem_simulation.cpp:204-325, 585-593)."""
import numpy as np
import pytest

from metakonga_amd import scenes, state_io
from metakonga_amd.batch import Batch
from metakonga_amd.xdem import Simulation
from test_gpu_drum import _drum_scene
from test_gpu_history import _run, _same, history_scene

pytestmark = pytest.mark.gpu


def _moving_drum_scene():
    sc, vel = _drum_scene(start=5e-5, end=1.5e-4)
    # a floor plane below the drum that rises from t = 3 dt until t = 1.2e-4
    sc.planes = [scenes.plane4((-0.1, -0.045, -0.1), (0.1, -0.045, -0.1), (0.1, -0.045, 0.1),
                               (-0.1, -0.045, 0.1))]
    sc.plane_motions = {0: scenes.motion(velocity=(0.0, 0.2, 0.0), start_time=3e-5,
                                         end_time=1.2e-4)}
    return sc, vel


def _sim_state(sim):
    st = sim.state()
    tris, sph = sim.polygons()
    objs = sim.objects()
    return st, bytes(tris), sph, bytes(objs), sim.steps


def test_sim_checkpoint_resumes_bitexact(lib, tmp_path):
    sc, vel = _moving_drum_scene()
    ref = Simulation(sc, lib)
    ref.set_state(vel=vel)
    ref.step(12)
    path = tmp_path / "sim.ckpt"
    ref.save_checkpoint(path)
    assert state_io.verify_checkpoint(path) == 1
    ref.step(18)           # crosses the end of both motions
    want = _sim_state(ref)
    ref.close()
    sim = Simulation(sc, lib)
    sim.load_checkpoint(path)
    assert sim.steps == 12
    sim.step(18)
    got = _sim_state(sim)
    sim.close()
    for k in want[0]:
        assert np.array_equal(got[0][k], want[0][k]), k
    assert got[1] == want[1] and np.array_equal(got[2], want[2]) and got[3] == want[3]
    assert got[4] == want[4] == 30


def test_sim_checkpoint_refuses_other_models(lib, tmp_path):
    sc, vel = _moving_drum_scene()
    sim = Simulation(sc, lib)
    sim.step(2)
    path = tmp_path / "sim.ckpt"
    sim.save_checkpoint(path)
    sim.close()
    other = scenes.Scene(name="o", pos=sc.pos[:-1].copy(), dt=sc.dt, nsteps=1,
                         world_origin=sc.world_origin, cell_size=0.004, grid_size=sc.grid_size)
    sim2 = Simulation(other, lib)
    with pytest.raises(RuntimeError):
        sim2.load_checkpoint(path)
    sim2.close()


def test_sim_part_final_and_bfr_resume(lib, tmp_path):
    sc, vel = _drum_scene()
    sim = Simulation(sc, lib)
    sim.set_state(vel=vel)
    sim.step(7)
    st = sim.state()
    sim.save_part(tmp_path, 7)
    ct, pos, prec = state_io.read_part(tmp_path / "part0007.bin")
    assert ct == 7 * sc.dt and prec == 2 and np.array_equal(pos, st["pos"])
    sim.save_final(tmp_path / "m_final.bfr")
    ct, P, V, A, prec = state_io.read_bfr(tmp_path / "m_final.bfr", sc.np)
    assert ct == 7 * sc.dt and np.array_equal(P, st["pos"]) and np.array_equal(V, st["vel"])
    assert np.array_equal(A, st["omega"])
    m, _ = sim.mass_inertia()
    sim.close()
    fresh = Simulation(sc, lib)
    fresh.load_bfr(tmp_path / "m_final.bfr")
    s2 = fresh.state()
    assert fresh.steps == 7
    assert np.array_equal(s2["pos"], st["pos"]) and np.array_equal(s2["vel"], st["vel"])
    assert np.array_equal(s2["omega"], st["omega"])
    assert not s2["acc"].any() and not s2["alpha"].any() and not s2["moment"].any()
    assert np.array_equal(s2["force"], m[:, None] * np.asarray(sc.gravity)[None, :])
    fresh.close()


def test_batch_checkpoint_resumes_bitexact(lib, tmp_path):
    h, vh, oh = history_scene(n_side=8, seed=41)
    plain = scenes.box_scene(1500, 3e-3, 77, 5e-6, 100)
    ref = _run([plain, h], 40, 20, vel=[None, vh], om=[None, oh])
    cap = plain.np + h.np
    ncell = int(np.prod(plain.grid_size)) + int(np.prod(h.grid_size))
    b = Batch(cap, 2, ncell)
    ids = [b.add(plain), b.add(h, vh, oh)]
    b.step(20)
    path = tmp_path / "batch.ckpt"
    b.save_checkpoint(path)
    b.close()
    assert state_io.verify_checkpoint(path) == 2
    b2 = Batch(cap, 2, ncell)
    ids2 = b2.load_checkpoint(path)
    assert len(ids2) == 2 and [b2.sizes[i] for i in ids2] == [plain.np, h.np]
    b2.step(20)
    out = [(b2.state(i), b2.history(i)) for i in ids2]
    b2.close()
    assert ref[1][1][0].sum() > 0
    _same(ref[0], out[0])
    _same(ref[1], out[1])
    assert ids
