"""The benchmark workloads at full size on the fused FP32c path.

* 1M-sphere hopper (radii U[2,4] mm, 6 planes, cell 8 mm, 128^3 grid): after
  300 steps from the seeded fill with random velocities of 1 m/s (so that
  spheres collide), the grid (sorted keys and ids, cell start/end) and the
  contact pair list bit-exact against the oracle on the same state, forces
  and torques of every particle within the FP32c bar (1e-5 of the
  particle's summed contact-law term magnitudes plus its weight; more than
  99.9% of the particles also within 1e-5 of their summed net contact-force
  magnitudes), and the state finite and inside the hopper.
* 256 box-settling scenes (N_s ~ logU[2k, 200k], 1.07e7 spheres) in one
  segmented grid, three 100-step graph blocks: three sampled scenes (the
  smallest, a median one, the largest) bit-identical to the same scene run
  alone.
* The serving loop (metakonga_amd.serve.Server) driving a real batch on a
  short seeded Poisson trace: every retired result bit-identical to the
  same scene run alone for the same steps; one batch's device-memory use
  matches the library's footprint figure.
* A NaN in one scene leaves the scenes sharing its batch bit-identical to
  their runs alone.
* The contact-rich compressed packing (200,000 spheres) with candidate lists:
  grid and pairs bit-exact, forces within the FP32c bar.
"""
import ctypes as C

import numpy as np
import pytest

from metakonga_amd import abi, scenes, serve
from metakonga_amd.batch import Batch
from xdem_testutil import FP32C_TOL, force_errors

pytestmark = pytest.mark.gpu
G = 9.80665


def _alone(sc, steps, block=100, vel=None):
    b = Batch(sc.np, 1, int(np.prod(sc.grid_size)))
    sid = b.add(sc, vel)
    done = 0
    while done < steps:
        b.step(min(block, steps - done))
        done += block
    st = b.state(sid)
    b.close()
    return st


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k)


def test_hopper_1m_grid_pairs_forces(lib, orc):
    sc = scenes.hopper()
    n = sc.np
    assert n == 1_000_000
    vel = scenes.as_f32_exact(scenes.uniform(0xABC, 3 * n, 5).reshape(n, 3) * 2.0 - 1.0)
    ncell = int(np.prod(sc.grid_size))
    b = Batch(n, 1, ncell)
    sid = b.add(sc, vel)
    for _ in range(3):
        b.step(100)
    st = b.state(sid)
    print(f"hopper 1M: after 300 steps, movers of the last sort {b.last_movers()}", flush=True)
    # finite, inside the hopper's side walls and above its bottom plates
    assert np.all(np.isfinite(st["pos"])) and np.all(np.isfinite(st["vel"]))
    hw = abs(sc.planes[0].xw.x)
    assert np.abs(st["pos"][:, 0]).max() < hw and np.abs(st["pos"][:, 2]).max() < hw
    assert st["pos"][:, 1].min() > 0.0
    b.prepare()
    out = b.probe(sid, ncell, max_per=32)
    b.close()
    # the oracle on the same state
    p = sc.params()
    pos = np.ascontiguousarray(st["pos"])
    h, idx = orc.hash_and_index(p, pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    assert np.array_equal(out["sorted_keys"], h)
    assert np.array_equal(out["sorted_ids"], srt)
    assert np.array_equal(out["cstart"], cs)
    assert np.array_equal(out["cend"], ce)
    m, I = sc.mass_inertia()
    f = np.zeros((n, 3))
    t = np.zeros((n, 3))
    vv = np.ascontiguousarray(st["vel"])
    om = np.ascontiguousarray(st["omega"])
    with orc.force_magnitudes(n) as fm:
        pairs = orc.p2p(p, 1, pos, vv, om, f, t, m, srt, cs, ce, sc.contact_pp(),
                        want_pairs=True, max_pairs=16 * n)
        for pl in sc.planes:
            orc.plane(p, 1, pl, pos, vv, om, f, t, m, sc.contact_pp())
    cnt = out["counts"]
    assert cnt.max() <= 32
    print(f"hopper 1M: {len(pairs)} contact pairs, {np.count_nonzero(fm)} particles in contact",
          flush=True)
    assert len(pairs) > 10_000
    got = np.repeat(np.arange(n, dtype=np.uint32), cnt)
    part = out["partners"][np.arange(32)[None, :] < cnt[:, None]]
    assert np.array_equal(np.stack([got, part], 1), pairs)
    fe = force_errors(out["force"], f, fm, m)
    te = force_errors(out["moment"], t, fm, m, pos[:, 3])
    print(f"hopper 1M: FP32c vs FP64 oracle force {fe}, torque {te}")
    assert fe["term"] < FP32C_TOL and te["term"] < FP32C_TOL
    assert fe["net_ok_share"] > 0.999


def test_dense_bed_grid_pairs_forces(lib, orc):
    """The contact-rich compressed packing (200,000 spheres, FCC at 0.985 x 2r,
    about 9 contacts each) after 100 steps with candidate lists: grid and
    contact pairs bit-exact against the oracle on the same state, forces and
    torques within the FP32c bar."""
    sc = scenes.dense_bed(n=200_000)
    n = sc.np
    ncell = int(np.prod(sc.grid_size))
    b = Batch(n, 1, ncell)
    b.lists(True, 0.05)
    sid = b.add(sc)
    b.step(50)
    b.step(50)
    st = b.state(sid)
    assert np.all(np.isfinite(st["pos"])) and np.all(np.isfinite(st["vel"]))
    b.prepare()
    out = b.probe(sid, ncell, max_per=32)
    b.close()
    p = sc.params()
    pos = np.ascontiguousarray(st["pos"])
    h, idx = orc.hash_and_index(p, pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    assert np.array_equal(out["sorted_ids"], srt)
    assert np.array_equal(out["cstart"], cs) and np.array_equal(out["cend"], ce)
    m, _ = sc.mass_inertia()
    f, t = np.zeros((n, 3)), np.zeros((n, 3))
    with orc.force_magnitudes(n) as fm:
        pairs = orc.p2p(p, 1, pos, np.ascontiguousarray(st["vel"]), np.ascontiguousarray(st["omega"]),
                        f, t, m, srt, cs, ce, sc.contact_pp(), want_pairs=True, max_pairs=16 * n)
        for pl in sc.planes:
            orc.plane(p, 1, pl, pos, np.ascontiguousarray(st["vel"]),
                      np.ascontiguousarray(st["omega"]), f, t, m, sc.contact_pp())
    cnt = out["counts"]
    print(f"dense bed: {len(pairs) / n:.2f} contacts per sphere", flush=True)
    assert len(pairs) > 8 * n
    got = np.repeat(np.arange(n, dtype=np.uint32), cnt)
    part = out["partners"][np.arange(32)[None, :] < cnt[:, None]]
    assert np.array_equal(np.stack([got, part], 1), pairs)
    fe = force_errors(out["force"], f, fm, m)
    te = force_errors(out["moment"], t, fm, m, pos[:, 3])
    print(f"dense bed: FP32c vs FP64 oracle force {fe}, torque {te}")
    assert fe["term"] < FP32C_TOL and te["term"] < FP32C_TOL


def test_batch_256_scenes_alone_vs_mix(lib):
    scs = scenes.batch_scenes()
    assert len(scs) == 256
    total = sum(s.np for s in scs)
    b = Batch(total, len(scs), sum(int(np.prod(s.grid_size)) for s in scs))
    ids = [b.add(s) for s in scs]
    for _ in range(3):
        b.step(100)
    sizes = np.array([s.np for s in scs])
    order = np.argsort(sizes)
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    mixed = {k: b.state(ids[k]) for k in picks}
    b.close()
    print(f"256 scenes, {total} spheres: comparing scenes {picks} (sizes {sizes[picks]})")
    for k in picks:
        _same(_alone(scs[k], 300), mixed[k], f"scene {k}")


def test_serving_results_equal_alone(lib):
    trace = serve.poisson_trace(seed=0xD3E4, rate=20.0, duration=1.5, n_lo=1000, n_hi=40_000,
                                steps_lo=200, steps_hi=600)
    assert len(trace) >= 15
    max_p, max_s, max_c = 200_000, 32, 1 << 25
    b = Batch(max_p, max_s, max_c)
    srv = serve.Server(b, max_particles=max_p, max_scenes=max_s, max_cells=max_c,
                       hbm_budget=serve.device_bytes(max_p, max_s, max_c), keep_results=True)
    done = srv.run(trace)
    b.close()
    served = [r for r in done if not (r.result or {}).get("rejected")]
    assert len(served) == len(trace)
    overlap = max(sum(1 for o in served if o.admitted < r.admitted < o.finished) for r in served)
    print(f"serving: {len(served)} requests, up to {overlap + 1} scenes in flight together")
    assert overlap >= 1  # continuous batching: requests ran together
    for r in served:
        _same(_alone(r.req.scene(), r.done_steps), r.result, f"request {r.req.rid}")


def test_batch_memory_matches_footprint(lib):
    free0, tot = C.c_size_t(), C.c_size_t()
    free1 = C.c_size_t()
    max_p, max_s, max_c = 8_000_000, 64, 1 << 26
    abi.check(lib, lib.xdem_synchronize())
    abi.check(lib, lib.xdem_device_memory(C.byref(free0), C.byref(tot)))
    b = Batch(max_p, max_s, max_c)
    sc = scenes.box_scene(20_000, 3e-3, 5, 5e-6, 100)
    b.add(sc)
    b.step(1)  # layout tables allocated
    b.sync()
    abi.check(lib, lib.xdem_device_memory(C.byref(free1), C.byref(tot)))
    used = free0.value - free1.value
    want = serve.device_bytes(max_p, max_s, max_c, lists=False)
    print(f"batch device memory {used / 1e6:.1f} MB, footprint {want / 1e6:.1f} MB")
    assert abs(used - want) <= 0.03 * want + (64 << 20), (used, want)
    # with candidate lists
    b.lists(True)
    b.step(1)
    b.sync()
    abi.check(lib, lib.xdem_device_memory(C.byref(free1), C.byref(tot)))
    used = free0.value - free1.value
    want = serve.device_bytes(max_p, max_s, max_c, lists=True)
    print(f"with lists {used / 1e6:.1f} MB, footprint {want / 1e6:.1f} MB")
    b.close()
    assert abs(used - want) <= 0.03 * want + (64 << 20), (used, want)


def test_nan_scene_quarantined(lib):
    """A NaN velocity in one scene: the other scenes of the batch stay
    bit-identical to their runs alone (each scene owns its cell keys, so
    nothing of the poisoned scene reaches their neighbour search)."""
    scs = [scenes.box_scene(n, r, 700 + k, 5e-6, 100) for k, (n, r) in
           enumerate([(3000, 3e-3), (2000, 4e-3), (4000, 2.5e-3)])]
    bad = np.zeros((scs[1].np, 3))
    bad[17, 1] = np.nan
    vels = [None, bad, None]
    b = Batch(sum(s.np for s in scs), 3, sum(int(np.prod(s.grid_size)) for s in scs))
    ids = [b.add(s, v) for s, v in zip(scs, vels)]
    b.step(50)
    b.step(50)
    mixed = [b.state(i) for i in ids]
    b.close()
    assert not np.all(np.isfinite(mixed[1]["vel"]))  # the poisoned scene carries its NaN
    for k in (0, 2):
        _same(_alone(scs[k], 100, block=50), mixed[k], f"scene {k}")
