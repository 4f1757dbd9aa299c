"""GPU parity of the fused FP32c batched path against the FP64 oracle.

Bars (stated per check):
  * cell keys, sorted ids, cell start/end and contact pair lists: bit-exact
    against the oracle on identical states (positions FP64, radii/velocities
    FP32-representable so both representations hold the same state);
  * forces and torques of one contact evaluation: within 1e-5 (FP32C_TOL) of
    the per-particle sum of the contact-law term magnitudes plus the weight
    (xdem_testutil: the oracle sums them exactly; FP32 arithmetic: ~3e-7
    measured); each test also prints the error against the sum of the net
    contact-force magnitudes, which exceeds the bar only where a contact's
    terms cancel;
  * a scene alone and inside a mix of scenes: bit-identical (each particle's
    arithmetic depends only on its own scene);
  * graph replay vs eager launches: bit-identical.
"""
import numpy as np
import pytest

from metakonga_amd import abi, scenes
from metakonga_amd.batch import Batch
from xdem_testutil import (FP32C_TOL, acc_error, dense_scene, force_errors, rel_error,
                           world_step_magnitudes)

pytestmark = pytest.mark.gpu

MAT = scenes.MaterialSpec()
CP = scenes.contact_property(MAT, MAT, scenes.ContactSpec())
G = 9.80665


def oracle_reference(orc, sc, vel, om):
    p = sc.params()
    n = sc.np
    h, idx = orc.hash_and_index(p, sc.pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    m, I = sc.mass_inertia()
    f = np.zeros((n, 3))
    t = np.zeros((n, 3))
    with orc.force_magnitudes(n) as fm:
        pairs = orc.p2p(p, 1, sc.pos, vel, om, f, t, m, srt, cs, ce, sc.contact_pp(),
                        want_pairs=True)
        for pl in sc.planes:
            orc.plane(p, 1, pl, sc.pos, vel, om, f, t, m, sc.contact_pp())
    return dict(h=h, srt=srt, cs=cs, ce=ce, pairs=pairs, f=f, t=t, m=m, fmag=fm.copy())


def probe_errors(out, ref, radius):
    """Force and torque errors of a probe against the oracle reference (the
    bar's term scale; the net-scale figures are printed by the callers)."""
    fe = force_errors(out["force"], ref["f"], ref["fmag"], ref["m"])
    te = force_errors(out["moment"], ref["t"], ref["fmag"], ref["m"], radius)
    print(f"  net scale: force {fe['net']:.2e} ({100 * fe['net_ok_share']:.3f}% within bar), "
          f"torque {te['net']:.2e}")
    return fe["term"], te["term"]


def test_probe_grid_pairs_forces(lib, orc):
    sc, vel, om = dense_scene()
    ref = oracle_reference(orc, sc, vel, om)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
    sid = b.add(sc, vel, om)
    b.prepare()
    out = b.probe(sid, int(np.prod(sc.grid_size)), max_per=32)
    # bit-exact grid
    assert np.array_equal(out["sorted_keys"], ref["h"])
    assert np.array_equal(out["sorted_ids"], ref["srt"])
    assert np.array_equal(out["cstart"], ref["cs"])
    assert np.array_equal(out["cend"], ref["ce"])
    # bit-exact contact pair list, traversal order per particle
    cnt = out["counts"]
    assert cnt.max() <= 32
    got = np.concatenate([np.stack([np.full(c, i, np.uint32), out["partners"][i, :c]], 1)
                          for i, c in enumerate(cnt) if c])
    assert len(got) > 4 * sc.np
    assert np.array_equal(got, ref["pairs"])
    # FP32 contact arithmetic
    ferr, terr = probe_errors(out, ref, sc.pos[:, 3])
    print(f"FP32c vs FP64 oracle: force err {ferr:.2e}, torque err {terr:.2e}")
    assert ferr < FP32C_TOL, ferr
    assert terr < FP32C_TOL, terr
    b.close()


def test_probe_cohesion(lib, orc):
    """Particle-particle cohesion on the fused path (coh pi R* delta, the
    reference's cohesionForce for a Hertz contact) against the oracle's
    cohesionForce (mphysics_cuda_impl.cuh:259-279), at the FP32c bar; the term
    is large enough to matter (1e5 Pa: about a third of the Hertz repulsion)."""
    sc, vel, om = dense_scene(n_side=10, seed=7)
    out = {}
    for coh in (0.0, 1e5):
        sc.contact = scenes.ContactSpec(cohesion=coh)
        ref = oracle_reference(orc, sc, vel, om)
        b = Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
        sid = b.add(sc, vel, om)
        b.prepare()
        out[coh] = b.probe(sid, int(np.prod(sc.grid_size)), max_per=32)
        b.close()
        ferr, terr = probe_errors(out[coh], ref, sc.pos[:, 3])
        print(f"cohesion {coh:g}: force err {ferr:.2e}, torque err {terr:.2e}")
        assert ferr < FP32C_TOL, (coh, ferr)
        assert terr < FP32C_TOL, (coh, terr)
    scale = ref["fmag"] + ref["m"] * G
    assert rel_error(out[1e5]["force"], out[0.0]["force"], scale) > 1e-2


def test_first_step_vs_oracle(lib, orc):
    """One full step from an identical state: velocities within the FP32c bound."""
    sc, vel, om = dense_scene(n_side=10, seed=5)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
    sid = b.add(sc, vel, om)
    b.graphs(False)
    b.step(1)
    st = b.state(sid)
    w = orc.World(sc)
    w.vel[:] = vel
    w.omega[:] = om
    fm = world_step_magnitudes(orc, w)
    assert np.abs(st["pos"] - w.pos).max() < 1e-9  # dt * FP32 velocity rounding
    # dv over one step = dt * F / m: compare m a = F per particle against the
    # step's contact-force magnitudes plus the weight
    err = acc_error(st["acc"], w, fm)
    print(f"first step: force err {err}")
    assert err["term"] < FP32C_TOL, err
    b.close()


def _run(scs, steps, block, graphs=True, vel=None, om=None):
    b = Batch(sum(s.np for s in scs), len(scs), sum(int(np.prod(s.grid_size)) for s in scs))
    b.graphs(graphs)
    ids = [b.add(s, None if vel is None else vel[k], None if om is None else om[k])
           for k, s in enumerate(scs)]
    done = 0
    while done < steps:
        b.step(min(block, steps - done))
        done += block
    out = [b.state(i) for i in ids]
    b.close()
    return out


def test_alone_vs_mix_bitexact(lib):
    scs = [scenes.box_scene(n, r, 100 + k, 5e-6, 100) for k, (n, r) in
           enumerate([(3000, 3e-3), (800, 4.5e-3), (5000, 2e-3)])]
    mixed = _run(scs, 60, 20)
    for k, s in enumerate(scs):
        alone = _run([s], 60, 20)[0]
        for key in alone:
            assert np.array_equal(alone[key], mixed[k][key]), (k, key)


def test_graph_replay_equals_eager(lib):
    sc, vel, om = dense_scene(n_side=9, seed=3)
    a = _run([sc], 75, 25, graphs=True, vel=[vel], om=[om])[0]
    e = _run([sc], 75, 25, graphs=False, vel=[vel], om=[om])[0]
    for key in a:
        assert np.array_equal(a[key], e[key]), key


@pytest.mark.parametrize("speed,side", [(0.05, 11), (600.0, 22), (600.0, 44)])
def test_incremental_sort_bitexact(lib, orc, speed, side):
    """After steps, the grid of the current positions equals the oracle's
    stable sort: few movers (the window patch), most of 10,648 particles
    changing cell every step (the movers ranked by counting across the blocks,
    more than the 2048 k_merge keeps in LDS), and most of 85,184 (more than
    the 32,768 ranked: the one-workgroup radix fallback)."""
    sc, vel, om = dense_scene(n_side=side, seed=9, pitch=2.3, jitter=0.1, vel_scale=speed)
    sc.planes = []
    sc.gravity = (0.0, 0.0, 0.0)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
    sid = b.add(sc, vel, om)
    b.step(7)
    b.step(5)
    moved = b.last_movers()
    if speed > 1:
        assert moved > (32768 if side > 30 else 2048), moved  # the large-M path was taken
    st = b.state(sid)
    b.prepare()
    out = b.probe(sid, int(np.prod(sc.grid_size)))
    p = sc.params()
    pos = np.ascontiguousarray(st["pos"])
    h, idx = orc.hash_and_index(p, pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    assert np.array_equal(out["sorted_keys"], h)
    assert np.array_equal(out["sorted_ids"], srt)
    assert np.array_equal(out["cstart"], cs)
    assert np.array_equal(out["cend"], ce)
    b.close()


def test_remove_and_add(lib):
    scs = [scenes.box_scene(n, r, 200 + k, 5e-6, 100) for k, (n, r) in
           enumerate([(2000, 3e-3), (1500, 4e-3), (2500, 2.5e-3), (1000, 3.5e-3)])]
    b = Batch(8000, 4, sum(int(np.prod(s.grid_size)) for s in scs))
    ids = [b.add(s) for s in scs[:3]]
    b.step(20)
    b.remove(ids[1])
    i3 = b.add(scs[3])
    b.step(20)
    b.step(20)
    s0 = b.state(ids[0])
    s2 = b.state(ids[2])
    s3 = b.state(i3)
    b.close()
    a0 = _run([scs[0]], 60, 20)[0]
    a2 = _run([scs[2]], 60, 20)[0]
    a3 = _run([scs[3]], 40, 20)[0]
    for k in a0:
        assert np.array_equal(a0[k], s0[k]), k
        assert np.array_equal(a2[k], s2[k]), k
        assert np.array_equal(a3[k], s3[k]), k


def test_capacity_is_enforced(lib):
    from metakonga_amd import abi
    sc = scenes.box_scene(1000, 3e-3, 1, 5e-6, 10)
    b = Batch(1500, 4, int(np.prod(sc.grid_size)) * 4)
    b.add(sc)
    with pytest.raises(abi.XdemError, match="status 4"):
        b.add(sc)
    b.close()


def test_energy_reduction(lib):
    sc, vel, om = dense_scene(n_side=8, seed=2)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    sid = b.add(sc, vel, om)
    b.step(1)
    st = b.state(sid)
    ke, mom = b.energy(sid)
    m = (2500.0 * 4 / 3 * np.pi) * st["pos"][:, 3].astype(np.float32).astype(np.float64) ** 3
    m = m.astype(np.float32).astype(np.float64)
    r = st["pos"][:, 3]
    I = 0.4 * m * r * r
    ke_ref = 0.5 * (m * (st["vel"] ** 2).sum(1)).sum() + 0.5 * (I * (st["omega"] ** 2).sum(1)).sum()
    assert ke == pytest.approx(ke_ref, rel=1e-6)
    assert np.allclose(mom, (m[:, None] * st["vel"]).sum(0), rtol=1e-6, atol=1e-12)
    b.close()


def test_checkpoint_resume_bitexact(lib):
    """State read back (with accelerations) and re-added resumes bit-exactly."""
    sc, vel, om = dense_scene(n_side=9, seed=13)
    ref = _run([sc], 40, 20, vel=[vel], om=[om])[0]
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    sid = b.add(sc, vel, om)
    b.step(20)
    ck = b.state(sid)
    b.close()
    b2 = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    sid2 = b2.add(sc, pos=ck["pos"], vel=ck["vel"], omega=ck["omega"], acc=ck["acc"],
                  alpha=ck["alpha"])
    b2.step(20)
    out = b2.state(sid2)
    b2.close()
    for k in ref:
        assert np.array_equal(ref[k], out[k]), k


def test_block_size_does_not_change_results(lib):
    """The autotuned block size only changes the launch shape, not the arithmetic."""
    sc, vel, om = dense_scene(n_side=9, seed=17)
    outs = []
    for bs in (64, 256):
        b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
        b.block(bs)
        sid = b.add(sc, vel, om)
        b.step(30)
        outs.append(b.state(sid))
        b.close()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    b.add(sc, vel, om)
    chosen, ms = b.autotune(steps_per_trial=5)
    assert chosen % 1000 in (64, 128, 256) and (chosen // 1000) % 10 in (1, 2, 4)
    assert (chosen // 10000) % 10 in (0, 1) and (chosen // 100000) % 10 in (0, 1)
    assert chosen // 1000000 in (0, 1)
    assert len(ms) == 54 and all(v > 0 for v in ms.values())
    b.close()


def test_capture_ahead_is_the_same_run(lib):
    """Graphs instantiated ahead (xdem_batch_capture, an odd block size too, so
    both buffer parities) replay exactly what the blocks captured at first use do."""
    sc, vel, om = dense_scene(n_side=9, seed=23)
    outs = []
    for ahead in (False, True):
        b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
        b.lists(True, 0.05)
        sid = b.add(sc, vel, om)
        b.prepare()
        if ahead:
            b.capture(7)
            b.capture(10)
        for k in (7, 7, 10, 7, 10):
            b.step(k)
        outs.append(b.state(sid))
        b.close()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_many_contacts_rescan(lib, orc):
    """A sphere touched by 40 smaller ones (more contacts than the 16-entry
    list): the further chunks are found by rescanning the cells; the pair list
    stays bit-exact and in traversal order, the forces within the FP32c bar."""
    R, r = float(scenes.as_f32_exact(5e-3)), float(scenes.as_f32_exact(1.2e-3))
    gold = np.pi * (3 - np.sqrt(5))
    n_small = 40
    k = np.arange(n_small)
    z = 1 - 2 * (k + 0.5) / n_small
    rr = np.sqrt(1 - z * z)
    dirs = np.c_[rr * np.cos(gold * k), rr * np.sin(gold * k), z]
    c = np.array([0.05, 0.05, 0.05])
    pos = np.vstack([np.r_[c, R], np.c_[c + dirs * (R + r - 2e-5), np.full(n_small, r)]])
    sc = scenes.Scene(name="crowd", pos=pos, dt=1e-5, nsteps=1, world_origin=(0.0, 0.0, 0.0),
                      cell_size=2 * R, grid_size=(16, 16, 16), gravity=(0.0, 0.0, 0.0))
    vel = scenes.as_f32_exact(scenes.uniform(7, 3 * sc.np).reshape(-1, 3) * 0.02 - 0.01)
    om = np.zeros_like(vel)
    ref = oracle_reference(orc, sc, vel, om)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
    sid = b.add(sc, vel, om)
    b.prepare()
    out = b.probe(sid, int(np.prod(sc.grid_size)), max_per=64)
    cnt = out["counts"]
    assert cnt[0] == n_small
    got = np.concatenate([np.stack([np.full(cc, i, np.uint32), out["partners"][i, :cc]], 1)
                          for i, cc in enumerate(cnt) if cc])
    assert np.array_equal(got, ref["pairs"])
    ferr, terr = probe_errors(out, ref, sc.pos[:, 3])
    print(f"many contacts: force err {ferr:.2e}, torque err {terr:.2e}")
    assert ferr < FP32C_TOL, ferr
    assert terr < FP32C_TOL, terr
    b.close()


def _variant_scenes(kind):
    if kind == "plain":
        sc, vel, om = dense_scene(n_side=9, seed=17)
        return sc, vel, om
    if kind == "history":
        from test_gpu_history import history_scene
        return history_scene(n_side=9, seed=23)
    from test_gpu_drum import _drum_scene
    sc, vel = _drum_scene()
    return sc, vel, None


@pytest.mark.parametrize("kind", ["plain", "history", "geometry"])
def test_walks_and_launch_shapes_bitexact(lib, kind):
    """Every contact-kernel variant (each lane walking its own rows or the wave
    walking row by row; 1, 2 or 4 candidates per load batch; contact records
    loaded ahead or not; block 64 or 256) visits the candidates in the same
    order and gives bit-identical states, on a dense box, with tangential
    history, and with a drum and a turning fin."""
    sc, vel, om = _variant_scenes(kind)
    ref = None
    for per_lane in (True, False):
        for unroll in (1, 2, 4):
            for ahead in (False, True):
                for block in ((64, 256) if unroll == 2 else (256,)):
                    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
                    b.launch(block, unroll, ahead)
                    b.walk(per_lane)
                    sid = b.add(sc, vel, om)
                    b.step(12)
                    st = b.state(sid)
                    b.close()
                    if ref is None:
                        ref = st
                        continue
                    for k in ref:
                        assert np.array_equal(ref[k], st[k]), (kind, per_lane, unroll, ahead,
                                                                block, k)
