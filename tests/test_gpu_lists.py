"""Candidate lists and the window-patch re-sort of the fused path against the
composition they replace (the 27-cell walk every step and the full merge
every step): bit-identical states after every run.

Lists keep, per particle, its neighbours within ri + rj + skin in the
reference traversal order; contacts found from a list are the walk's
contacts in the walk's order, so every result must match bit for bit —
across skins, launch shapes, graph replay, scene mixes, step-block lengths
(the lists carry over block boundaries), few movers (the window patch) and
many (the full merge and a rebuild of every list).
"""
import numpy as np
import pytest

from metakonga_amd import scenes
from metakonga_amd.batch import Batch
from xdem_testutil import dense_scene

pytestmark = pytest.mark.gpu


def _scenes(kind):
    if kind == "dense":  # contact-rich, slow: few particles change cell
        sc, vel, om = dense_scene(n_side=10, seed=21, vel_scale=0.05)
        return [sc], [vel], [om]
    if kind == "fast":  # most particles change cell every few steps
        sc, vel, om = dense_scene(n_side=12, seed=22, pitch=2.2, jitter=0.1, vel_scale=0.8)
        sc.gravity = (0.0, 0.0, 0.0)
        return [sc], [vel], [om]
    if kind == "boxes":  # three settling boxes in one batch
        scs = [scenes.box_scene(n, r, 500 + k, 5e-6, 100) for k, (n, r) in
               enumerate([(3000, 3e-3), (1200, 4.5e-3), (4000, 2e-3)])]
        return scs, [None] * 3, [None] * 3
    if kind == "hopper":  # polydisperse, colliding at 0.3 m/s
        sc = scenes.hopper(n=60_000)
        n = sc.np
        vel = scenes.as_f32_exact(scenes.uniform(0x51, 3 * n, 5).reshape(n, 3) * 0.6 - 0.3)
        return [sc], [vel], [None]
    if kind == "hopper_mid":  # a few hundred movers per step: the full merge, movers ranked
        sc = scenes.hopper(n=200_000)
        n = sc.np
        vel = scenes.as_f32_exact(scenes.uniform(0x52, 3 * n, 5).reshape(n, 3) * 1.0 - 0.5)
        return [sc], [vel], [None]
    if kind == "bed":  # compressed FCC packing: about 9 contacts per sphere
        sc = scenes.dense_bed(n=40_000)
        return [sc], [None], [None]
    if kind == "crowd":  # 24 contacts on one sphere: more than the 16-entry contact list
        R, r = float(scenes.as_f32_exact(5e-3)), float(scenes.as_f32_exact(1.2e-3))
        k = np.arange(24)
        z = 1 - 2 * (k + 0.5) / 24
        rr = np.sqrt(1 - z * z)
        g = np.pi * (3 - np.sqrt(5))
        dirs = np.c_[rr * np.cos(g * k), rr * np.sin(g * k), z]
        c = np.array([0.05, 0.05, 0.05])
        pos = np.vstack([np.r_[c, R], np.c_[c + dirs * (R + r - 2e-5), np.full(24, r)]])
        sc = scenes.Scene(name="crowd", pos=pos, dt=1e-5, nsteps=1, world_origin=(0.0, 0.0, 0.0),
                          cell_size=2 * R, grid_size=(16, 16, 16), gravity=(0.0, 0.0, 0.0))
        vel = scenes.as_f32_exact(scenes.uniform(9, 3 * sc.np).reshape(-1, 3) * 0.02 - 0.01)
        return [sc], [vel], [None]
    raise ValueError(kind)


def _run(kind, steps=90, block=30, patch=True, lists=False, skin=0.0, threads=256, unroll=2,
         ahead=False, graphs=True):
    scs, vels, oms = _scenes(kind)
    b = Batch(sum(s.np for s in scs), len(scs), sum(int(np.prod(s.grid_size)) for s in scs))
    b.graphs(graphs)
    b.launch(threads, unroll, ahead)
    b.sort_patch(patch)
    if lists:
        b.lists(True, skin)
    ids = [b.add(s, v, o) for s, v, o in zip(scs, vels, oms)]
    done, rebuilds, movers = 0, [], []
    while done < steps:
        b.step(min(block, steps - done))
        done += block
        rebuilds.append(b.list_rebuilds())
        movers.append(b.last_movers())
    out = [b.state(i) for i in ids]
    b.close()
    return out, rebuilds, movers


VARIANTS = [
    dict(patch=True),
    dict(patch=True, lists=True, skin=0.05),
    dict(patch=True, lists=True, skin=0.01),
    dict(patch=True, lists=True, skin=0.3),
    dict(patch=True, lists=True, skin=0.05, threads=64, unroll=4),
    dict(patch=True, lists=True, skin=0.05, unroll=1, ahead=True),
    dict(patch=True, lists=True, skin=0.05, graphs=False),
    # lists kept across many block boundaries (a block's last step does not
    # advance; the next block's first position update records its displacement)
    dict(patch=True, lists=True, skin=0.05, block=7),
    dict(patch=True, lists=True, skin=0.02, block=1, graphs=False),
]


@pytest.mark.parametrize("kind", ["dense", "fast", "boxes", "hopper", "hopper_mid", "bed", "crowd"])
def test_lists_and_patch_bitexact(lib, kind):
    ref, _, ref_movers = _run(kind, patch=False)
    print(f"{kind}: movers of the last sort of each block {ref_movers}")
    for v in VARIANTS:
        out, rebuilds, _ = _run(kind, **v)
        for k, (a, b) in enumerate(zip(ref, out)):
            for key in a:
                assert np.array_equal(a[key], b[key]), (kind, v, k, key)
        if kind == "hopper_mid":
            # more than 64 movers per sort (the full merge; the movers sorted by
            # counting ranks across the blocks)
            assert 64 < max(ref_movers) <= 32768, ref_movers
        if v.get("lists") and kind in ("dense", "bed", "crowd"):
            # quiet scenes: the lists were reused (no scene rebuilt them all)
            assert rebuilds[-1] == 0, (kind, v, rebuilds)


def test_lists_alone_vs_mix(lib):
    boxes, _, _ = _scenes("boxes")
    dense, dv, do = _scenes("dense")
    scs = [boxes[0], dense[0], boxes[2]]
    vels, oms = [None, dv[0], None], [None, do[0], None]

    def run(sel):
        b = Batch(sum(scs[k].np for k in sel), len(sel),
                  sum(int(np.prod(scs[k].grid_size)) for k in sel))
        b.lists(True, 0.1)
        ids = [b.add(scs[k], vels[k], oms[k]) for k in sel]
        for _ in range(3):
            b.step(25)
        out = [b.state(i) for i in ids]
        b.close()
        return out

    mixed = run([0, 1, 2])
    for k in range(3):
        alone = run([k])[0]
        for key in alone:
            assert np.array_equal(alone[key], mixed[k][key]), (k, key)


def test_batch_many_movers_window_patch(lib):
    """The 256 box scenes (1.07e7 spheres) over 2,400 steps: the particles that
    change cell per step grow past 512 (the block path sorting them with its
    largest networks), the window patch keeps every list (no scene rebuilds
    them all), and three sampled scenes end bit-identical to the run that takes
    the full merge every step."""
    scs = scenes.batch_scenes()
    ntot = sum(s.np for s in scs)
    cells = sum(int(np.prod(s.grid_size)) for s in scs)
    picks = [0, 128, 255]
    out = []
    for patch in (True, False):
        b = Batch(ntot, len(scs), cells)
        b.lists(True, 0.05)
        b.sort_patch(patch)
        ids = [b.add(s) for s in scs]
        trace = []
        for _ in range(24):
            b.step(100)
            trace.append((b.last_movers(), b.list_rebuilds()))
        out.append([b.state(ids[k]) for k in picks])
        b.close()
        if patch:
            print(f"movers and scenes rebuilding, per 100-step block: {trace}")
            assert any(512 < m <= 1024 and r == 0 for m, r in trace), trace
    for k in range(len(picks)):
        for key in out[0][k]:
            assert np.array_equal(out[0][k][key], out[1][k][key]), (picks[k], key)


def test_lists_carry_over_step_calls(lib):
    """A quiet scene keeps its lists from one step call to the next (the
    block's last step keeps them, the next block's position update records
    its displacement): no scene rebuilds every list at a block start, and the
    result equals one long call and the walk."""
    sc, vel, om = dense_scene(n_side=10, seed=21, vel_scale=0.05)

    def run(lists, blocks):
        b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
        if lists:
            b.lists(True, 0.05)
        sid = b.add(sc, vel, om)
        rebuilds = []
        for k in blocks:
            b.step(k)
            rebuilds.append(b.list_rebuilds())
        out = b.state(sid)
        b.close()
        return out, rebuilds

    ref, _ = run(False, [40])
    one, _ = run(True, [40])
    many, rb = run(True, [10, 1, 1, 1, 7, 20])
    for out in (one, many):
        for key in ref:
            assert np.array_equal(ref[key], out[key]), key
    # the first call builds the lists; every later block start reuses them
    assert rb[1:] == [0] * (len(rb) - 1), rb
