"""GPU parity of the tangential-spring history (fused FP32c path) against the
FP64 oracle.

Teacher-forced steps: after every GPU step the oracle restarts from the GPU's
own state and history and runs one step; the two must agree
  * bit-exactly on the history keys (per particle: partner ids in traversal
    order, then touching planes in order) and their counts — positions are
    FP64 and the FP32 velocities/accelerations are identical in both;
  * within FP32 rounding on the springs xi (HIST_TOL, relative to the
    particle's largest spring) and on the forces m a (FP32C_TOL = 1e-5,
    relative to the particle's summed contact-force magnitudes plus weight).
Also: a history scene alone and in a mix with memoryless scenes, graph replay
vs eager, checkpoint/resume with history — all bit-identical; slot overflow
counted as the oracle counts it.  Drums carry springs too (shell and end caps,
after the planes); the same teacher-forced comparison covers a history scene
inside a spinning drum, and such a scene alone and in a mix with a drum and
mesh scene without history is bit-identical.
"""
import numpy as np
import pytest

from metakonga_amd import scenes
from metakonga_amd.batch import Batch
from xdem_testutil import FP32C_TOL, acc_error, dense_scene, drum_history_scene

pytestmark = pytest.mark.gpu

HIST_TOL = 1e-4


def f32_mass(sc):
    r = sc.pos[:, 3].astype(np.float32)
    mc = np.float32(sc.material.density * (4.0 / 3.0) * np.pi)
    m = (mc * (r * r * r)).astype(np.float32).astype(np.float64)
    return m, 0.4 * m * sc.pos[:, 3] ** 2


def oracle_step(orc, sc, st, hist):
    """One oracle step from a GPU state + history."""
    m, I = f32_mass(sc)
    w = orc.World(sc, mass=m, iner=I)
    for k in ("vel", "omega", "acc", "alpha"):
        getattr(w, k)[:] = st[k]
    w.pos[:] = st["pos"]
    w.force[:] = m[:, None] * np.asarray(sc.gravity)[None, :]
    w.moment[:] = 0
    n = sc.np
    hin, hout = orc.History(n), orc.History(n)
    hin.count[:], hin.key[:], hin.xi[:] = hist
    with orc.force_magnitudes(n) as fm:
        w.hist_step(hin, hout)
    w.fmag = fm.copy()
    return w, hout


def compare(orc, sc, st, hist, w, hout):
    cnt, key, xi = hist
    assert np.array_equal(cnt, hout.count)
    live = np.arange(key.shape[1])[None, :] < cnt[:, None]
    assert np.array_equal(np.where(live, key, 0), np.where(live, hout.key, 0))
    ref = np.where(live[..., None], hout.xi, 0.0)
    got = np.where(live[..., None], xi, 0.0)
    scale = np.abs(ref).max(axis=(1, 2)) + 1e-30
    err = (np.abs(got - ref).max(axis=(1, 2)) / scale).max() if len(scale) else 0.0
    assert err < HIST_TOL, err
    ae = acc_error(st["acc"], w, w.fmag)
    aerr = ae["term"]
    print(f"history step: spring err {err:.2e}, force err {ae}")
    assert aerr < FP32C_TOL, ae
    return err, aerr


def history_scene(**kw):
    sc, vel, om = dense_scene(**kw)
    sc.tangential_history = True
    return sc, vel, om


@pytest.mark.parametrize("coh", [0.0, 1e5])
def test_history_steps_vs_oracle(lib, orc, coh):
    """coh > 0: the particle pairs' cohesionForce term with the springs."""
    sc, vel, om = history_scene(n_side=9, seed=31)
    sc.contact = scenes.ContactSpec(cohesion=coh)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    b.graphs(False)
    sid = b.add(sc, vel, om)
    prev = (b.state(sid), b.history(sid))
    assert prev[1][0].sum() == 0
    kept = 0
    for k in range(12):
        b.step(1)
        st, hist = b.state(sid), b.history(sid)
        w, hout = oracle_step(orc, sc, prev[0], prev[1])
        assert np.array_equal(st["pos"], w.pos)
        compare(orc, sc, st, hist, w, hout)
        # springs carried from the previous step (same key present twice)
        for i in range(sc.np):
            c0, c1 = int(prev[1][0][i]), int(hist[0][i])
            kept += len(set(prev[1][1][i, :c0]) & set(hist[1][i, :c1]))
        prev = (st, hist)
    assert kept > 12 * sc.np  # most contacts persist from step to step
    assert b.history_overflow() == 0
    walls = sum(int((prev[1][1][i, :prev[1][0][i]] >= 0x80000000).sum()) for i in range(sc.np))
    assert walls > 0  # plane springs are part of the history
    b.close()


def test_history_survives_resorts(lib, orc):
    """Particles that change cell keep their springs (history follows the
    particle through the incremental re-sort)."""
    sc, vel, om = history_scene(n_side=10, seed=33, pitch=1.99, jitter=0.1, vel_scale=0.6)
    sc.gravity = (0.0, 0.0, 0.0)
    # lattice sites on cell faces (not mid-cell), so that particles change cell
    sc.world_origin = tuple(np.asarray(sc.world_origin) + sc.pos[0, 3])
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    b.graphs(False)
    sid = b.add(sc, vel, om)
    b.step(1)
    prev = (b.state(sid), b.history(sid))
    p = sc.params()
    moved_kept = 0
    movers = 0
    for k in range(10):
        b.step(1)
        movers += b.last_movers()
        st, hist = b.state(sid), b.history(sid)
        w, hout = oracle_step(orc, sc, prev[0], prev[1])
        compare(orc, sc, st, hist, w, hout)
        h0, _ = orc.hash_and_index(p, np.ascontiguousarray(prev[0]["pos"]))
        h1, _ = orc.hash_and_index(p, np.ascontiguousarray(st["pos"]))
        for i in np.nonzero(h0 != h1)[0]:
            c0, c1 = int(prev[1][0][i]), int(hist[0][i])
            moved_kept += len(set(prev[1][1][i, :c0]) & set(hist[1][i, :c1]))
        prev = (st, hist)
    assert movers > 0 and moved_kept > 0, (movers, moved_kept)
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
    out = [(b.state(i), b.history(i)) for i in ids]
    b.close()
    return out


def _same(a, b):
    for k in a[0]:
        assert np.array_equal(a[0][k], b[0][k]), k
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)


def test_history_alone_vs_mix_and_graphs(lib):
    h, vh, oh = history_scene(n_side=8, seed=35)
    plain = [scenes.box_scene(n, r, 300 + k, 5e-6, 100) for k, (n, r) in
             enumerate([(2000, 3e-3), (1200, 4e-3)])]
    mixed = _run([plain[0], h, plain[1]], 60, 20, vel=[None, vh, None], om=[None, oh, None])
    _same(_run([h], 60, 20, vel=[vh], om=[oh])[0], mixed[1])
    _same(_run([h], 60, 20, graphs=False, vel=[vh], om=[oh])[0], mixed[1])
    for k, s in ((0, plain[0]), (2, plain[1])):
        alone = _run([s], 60, 20)[0]
        _same(alone, mixed[k])
        assert mixed[k][1][0].sum() == 0


@pytest.mark.parametrize("drum", [False, True])
def test_history_checkpoint_resume(lib, drum):
    sc, vel, om = (drum_history_scene if drum else history_scene)(n_side=8, seed=37)
    ref = _run([sc], 40, 20, vel=[vel], om=[om])[0]
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    sid = b.add(sc, vel, om)
    b.step(20)
    st, hist = b.state(sid), b.history(sid)
    b.close()
    assert hist[0].sum() > 0
    b2 = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    sid2 = b2.add(sc, pos=st["pos"], vel=st["vel"], omega=st["omega"], acc=st["acc"],
                  alpha=st["alpha"], history=hist)
    b2.step(20)
    out = (b2.state(sid2), b2.history(sid2))
    b2.close()
    _same(ref, out)
    if drum:
        assert (hist[1] >= 0xC0000000).any()
        bad = hist[1].copy()
        i = int(np.argmax(hist[0] > 0))
        bad[i, 0] = 0xC0000003            # no drum surface has side 3
        b3 = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
        with pytest.raises(RuntimeError):
            b3.add(sc, pos=st["pos"], vel=st["vel"], omega=st["omega"], acc=st["acc"],
                   alpha=st["alpha"], history=(hist[0], bad, hist[2]))
        b3.close()


def test_history_overflow_matches_oracle(lib, orc):
    """A large sphere touched by more small spheres than there are slots: the
    first 16 contacts keep springs, the rest are counted."""
    R, r = float(scenes.as_f32_exact(5e-3)), float(scenes.as_f32_exact(1.2e-3))
    dirs = []
    gold = np.pi * (3 - np.sqrt(5))
    n_small = 40
    for k in range(n_small):
        z = 1 - 2 * (k + 0.5) / n_small
        rr = np.sqrt(1 - z * z)
        dirs.append([rr * np.cos(gold * k), rr * np.sin(gold * k), z])
    dirs = np.array(dirs)
    c = np.array([0.05, 0.05, 0.05])
    small = c + dirs * (R + r - 2e-5)
    pos = np.vstack([np.r_[c, R], np.c_[small, np.full(n_small, r)]])
    sc = scenes.Scene(name="crowd", pos=pos, dt=1e-5, nsteps=1, world_origin=(0.0, 0.0, 0.0),
                      cell_size=2 * R, grid_size=(16, 16, 16), gravity=(0.0, 0.0, 0.0),
                      tangential_history=True)
    vel = scenes.as_f32_exact(scenes.uniform(5, 3 * sc.np).reshape(-1, 3) * 0.02 - 0.01)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    sid = b.add(sc, vel)
    st0 = b.state(sid)
    b.step(1)
    st, hist = b.state(sid), b.history(sid)
    w, hout = oracle_step(orc, sc, st0, (np.zeros(sc.np, np.uint32),
                                         np.zeros((sc.np, 16), np.uint32),
                                         np.zeros((sc.np, 16, 3))))
    compare(orc, sc, st, hist, w, hout)
    assert hout.overflow > 0
    assert b.history_overflow() == hout.overflow
    b.close()


@pytest.mark.parametrize("drum", [False, True])
def test_history_walks_and_blocks_bitexact(lib, drum):
    """Tangential history under both neighbour walks and block sizes 64, 128
    and 256: identical states, history counts, keys and springs.  The per-lane
    walk keeps its rows past the second in the thread's own history-key slots
    (rows in slots 2i, 2i+1); in a dense lattice (pitch just under the
    diameter, cell size = the diameter) most particles have more than two
    non-empty neighbour rows, so that table is used on nearly every lane
    while other waves of the block write their keys.  With drum: the same
    lattice inside a spinning drum (the history and drum variant)."""
    sc, vel, om = (drum_history_scene if drum else history_scene)(n_side=11, seed=41,
                                                                 jitter=0.08)
    ref = None
    for per_lane in (True, False):
        for block in (64, 128, 256):
            b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
            b.launch(block, 2, False)
            b.walk(per_lane)
            sid = b.add(sc, vel, om)
            b.step(15)
            out = (b.state(sid), b.history(sid))
            b.close()
            assert out[1][0].sum() > 4 * sc.np
            if ref is None:
                ref = out
                continue
            _same(ref, out)


def test_history_drum_steps_vs_oracle(lib, orc):
    """Teacher-forced steps of a history scene inside a spinning drum (caps on
    the lattice's top and bottom layers, shell on its vertical edges) and an
    open box of planes: keys bit-exact, springs and forces within FP32
    rounding of the oracle."""
    sc, vel, om = drum_history_scene(n_side=9, seed=53)
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    b.graphs(False)
    sid = b.add(sc, vel, om)
    prev = (b.state(sid), b.history(sid))
    kept, sides = 0, set()
    for k in range(12):
        b.step(1)
        st, hist = b.state(sid), b.history(sid)
        w, hout = oracle_step(orc, sc, prev[0], prev[1])
        assert np.array_equal(st["pos"], w.pos)
        compare(orc, sc, st, hist, w, hout)
        for i in range(sc.np):
            c0, c1 = int(prev[1][0][i]), int(hist[0][i])
            now = hist[1][i, :c1]
            kept += len({x for x in prev[1][1][i, :c0] if x >= 0xC0000000} & set(now))
            sides |= {int(x) & 3 for x in now if x >= 0xC0000000}
        prev = (st, hist)
    assert sides == {0, 1, 2}      # shell and both caps
    assert kept > 12 * 20          # drum springs carried from step to step
    assert b.history_overflow() == 0
    b.close()


def test_history_drum_alone_vs_mix_and_graphs(lib):
    from test_gpu_drum import _drum_scene
    h, vh, oh = drum_history_scene(n_side=8, seed=55)
    d, vd = _drum_scene()
    mixed = _run([d, h], 60, 20, vel=[vd, vh], om=[None, oh])
    _same(_run([h], 60, 20, vel=[vh], om=[oh])[0], mixed[1])
    _same(_run([h], 60, 20, graphs=False, vel=[vh], om=[oh])[0], mixed[1])
    _same(_run([d], 60, 20, vel=[vd])[0], mixed[0])
    assert mixed[0][1][0].sum() == 0 and mixed[1][1][0].sum() > 0
