"""ctypes wrapper of the fused FP32c batched path (include/xdem_batch.h).

Every computation runs in libxdem.so; this module only marshals scenes
(metakonga_amd.scenes.Scene) in and states out.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import abi

PHASES = ["pos_hash", "contact_verlet", "sort_patch", "sort_movers", "sort_merge", "cell_ranges"]


class BatchSceneDesc(C.Structure):
    _fields_ = [("np", C.c_uint32), ("dt", C.c_double), ("gravity", abi.double3),
                ("world_origin", abi.double3), ("cell_size", C.c_double), ("grid_size", abi.uint3),
                ("rolling", C.c_bool), ("particle", abi.Material), ("pp", abi.ContactPairDesc),
                ("wall", abi.ContactPairDesc), ("nplane", C.c_uint32), ("planes", C.c_void_p),
                ("tangential_history", C.c_int32)]


HISTORY_SLOTS = 16
WALL_KEY = 0x80000000
DRUM_KEY = 0xC0000000


def declare(lib):
    vp, u32, i32, u64 = C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64
    sig = {
        "xdem_batch_create": ([u32, u32, u64, C.POINTER(vp)], i32),
        "xdem_batch_device_bytes": ([u32, u32, u64, i32], u64),
        "xdem_batch_destroy": ([vp], i32),
        "xdem_batch_add_scene": ([vp, C.POINTER(BatchSceneDesc), vp, vp, vp, C.POINTER(u32)], i32),
        "xdem_batch_add_scene_state": ([vp, C.POINTER(BatchSceneDesc), vp, vp, vp, vp, vp,
                                        C.POINTER(u32)], i32),
        "xdem_batch_add_scene_history": ([vp, C.POINTER(BatchSceneDesc), vp, vp, vp, vp, vp, vp,
                                          vp, vp, C.POINTER(u32)], i32),
        "xdem_batch_get_history": ([vp, u32, vp, vp, vp], i32),
        "xdem_batch_scene_particles": ([vp, u32], u32),
        "xdem_batch_add_drum": ([vp, u32, vp, vp], i32),
        "xdem_batch_add_polygon_object": ([vp, u32, vp, u32, vp, vp, vp, vp], i32),
        "xdem_batch_save_checkpoint": ([vp, C.c_char_p], i32),
        "xdem_batch_load_checkpoint": ([vp, C.c_char_p, vp, u32, vp], i32),
        "xdem_batch_history_overflow": ([vp, C.POINTER(u32)], i32),
        "xdem_batch_remove_scene": ([vp, u32], i32),
        "xdem_batch_step": ([vp, u32], i32),
        "xdem_batch_set_graphs": ([vp, i32], i32),
        "xdem_batch_set_block": ([vp, u32], i32),
        "xdem_batch_set_launch": ([vp, u32, u32, i32], i32),
        "xdem_batch_set_walk": ([vp, i32], i32),
        "xdem_batch_set_lists": ([vp, i32, C.c_float], i32),
        "xdem_batch_list_rebuilds": ([vp, C.POINTER(u32)], i32),
        "xdem_batch_set_sort_patch": ([vp, i32], i32),
        "xdem_batch_synchronize": ([vp], i32),
        "xdem_batch_autotune": ([vp, u32, C.POINTER(u32), vp], i32),
        "xdem_batch_get_state": ([vp, u32, vp, vp, vp, vp, vp], i32),
        "xdem_batch_num_particles": ([vp], u32),
        "xdem_batch_num_scenes": ([vp], u32),
        "xdem_batch_energy": ([vp, u32, vp, vp], i32),
        "xdem_batch_enable_reactions": ([vp, i32], i32),
        "xdem_batch_object_reactions": ([vp, u32, vp, vp, u32, vp], i32),
        "xdem_batch_drum_reactions": ([vp, u32, vp, vp, u32, vp], i32),
        "xdem_batch_enable_timing": ([vp, i32], i32),
        "xdem_batch_phase_times": ([vp, vp, vp], i32),
        "xdem_batch_last_movers": ([vp, C.POINTER(u32)], i32),
        "xdem_batch_prepare": ([vp], i32),
        "xdem_batch_capture": ([vp, u32], i32),
        "xdem_batch_contact_metrics": ([vp, u32, C.POINTER(C.c_uint64), C.POINTER(C.c_double)], i32),
        "xdem_batch_probe": ([vp, u32, vp, vp, u32, vp, vp, vp, vp, vp, vp], i32),
    }
    for n, (a, r) in sig.items():
        f = getattr(lib, n, None)
        if f is not None:
            f.argtypes = a
            f.restype = r


def available(lib) -> bool:
    return hasattr(lib, "xdem_batch_create")


def scene_desc(scene):
    m = scene.material
    mat = abi.Material(m.density, m.youngs, m.poisson, m.shear)
    c = scene.contact
    d = BatchSceneDesc()
    d.np = scene.np
    d.dt = scene.dt
    d.gravity = abi._d3(scene.gravity)
    d.world_origin = abi._d3(scene.world_origin)
    d.cell_size = scene.cell_size or scene.default_cell_size()
    d.grid_size = abi.uint3(*scene.grid_size)
    d.rolling = scene.rolling
    d.particle = mat
    d.pp = abi.ContactPairDesc(c.restitution, c.friction, c.rolling_friction, c.cohesion,
                               c.stiffness_ratio, mat)
    d.wall = abi.ContactPairDesc(c.restitution, c.friction, c.rolling_friction, c.cohesion,
                                 c.stiffness_ratio, mat)
    planes = abi.struct_array(abi.PlaneInfo, scene.planes) if scene.planes else None
    d.nplane = len(scene.planes)
    d.planes = C.cast(planes, C.c_void_p) if planes is not None else None
    d.tangential_history = int(bool(getattr(scene, "tangential_history", False)))
    return d, planes


class Batch:
    """Scenes advanced together on the fused FP32c path."""

    def __init__(self, max_particles, max_scenes=1, max_cells=1 << 22, lib=None):
        self.lib = lib or abi.load()
        declare(self.lib)
        h = C.c_void_p()
        abi.check(self.lib, self.lib.xdem_batch_create(int(max_particles), int(max_scenes),
                                                       int(max_cells), C.byref(h)))
        self.h = h
        self.sizes = {}

    def add(self, scene, vel=None, omega=None, acc=None, alpha=None, pos=None, history=None):
        """Add a scene; pos/vel/omega/acc/alpha (optional) override the scene's
        initial state, `history` = (count, key, xi) as returned by history()
        restores the tangential springs (a checkpoint read back with state()
        and history() resumes bit-exactly)."""
        d, keep = scene_desc(scene)
        sid = C.c_uint32()
        pos = abi.as_f64(scene.pos if pos is None else pos)
        arr = [None if x is None else abi.as_f64(x) for x in (vel, omega, acc, alpha)]
        if history is None:
            hist = (None, None, None)
        else:
            hist = (np.ascontiguousarray(history[0], np.uint32),
                    np.ascontiguousarray(history[1], np.uint32), abi.as_f64(history[2]))
        abi.check(self.lib, self.lib.xdem_batch_add_scene_history(
            self.h, C.byref(d), abi.ptr(pos), *(abi.ptr(x) for x in arr),
            *(abi.ptr(x) for x in hist), C.byref(sid)))
        del keep
        for dr in scene.drums:
            abi.check(self.lib, self.lib.xdem_batch_add_drum(self.h, sid.value, C.byref(dr),
                                                             C.byref(d.wall)))
        if scene.tri is not None:
            dpi, _, dpmi = scene.polygon_arrays()
            ids = np.array([t.id for t in dpi])
            objs = scene.objects or []
            for o in range(len(dpmi)):
                sel = [dpi[k] for k in np.nonzero(ids == o)[0]]
                if not sel:
                    continue
                tris = abi.struct_array(abi.PolygonInfo, sel)
                mot = objs[o].motion if o < len(objs) else None
                abi.check(self.lib, self.lib.xdem_batch_add_polygon_object(
                    self.h, sid.value, C.cast(tris, C.c_void_p), len(sel),
                    C.byref(dpmi[o].origin), C.byref(dpmi[o].ep),
                    C.byref(mot) if mot is not None else None, C.byref(d.wall)))
        self.sizes[sid.value] = scene.np
        return sid.value

    def save_checkpoint(self, path):
        """Every scene's description, state and tangential history to a file."""
        abi.check(self.lib, self.lib.xdem_batch_save_checkpoint(self.h, os.fsencode(path)))

    def load_checkpoint(self, path, max_scenes=1 << 16):
        """Adds a checkpoint's scenes; returns their new ids in checkpoint order."""
        ids = np.zeros(max_scenes, np.uint32)
        n = C.c_uint32()
        abi.check(self.lib, self.lib.xdem_batch_load_checkpoint(
            self.h, os.fsencode(path), abi.ptr(ids), max_scenes, C.byref(n)))
        out = [int(x) for x in ids[:n.value]]
        for sid in out:
            self.sizes[sid] = int(self.lib.xdem_batch_scene_particles(self.h, sid))
        return out

    def remove(self, sid):
        abi.check(self.lib, self.lib.xdem_batch_remove_scene(self.h, sid))
        self.sizes.pop(sid, None)

    def step(self, n=1):
        abi.check(self.lib, self.lib.xdem_batch_step(self.h, int(n)))

    def sync(self):
        abi.check(self.lib, self.lib.xdem_batch_synchronize(self.h))

    def graphs(self, on: bool):
        abi.check(self.lib, self.lib.xdem_batch_set_graphs(self.h, int(on)))

    def block(self, threads: int):
        abi.check(self.lib, self.lib.xdem_batch_set_block(self.h, int(threads)))

    def launch(self, threads=256, unroll=2, load_ahead=False):
        abi.check(self.lib, self.lib.xdem_batch_set_launch(self.h, int(threads), int(unroll),
                                                           int(bool(load_ahead))))

    def walk(self, per_lane=True):
        abi.check(self.lib, self.lib.xdem_batch_set_walk(self.h, int(bool(per_lane))))

    def lists(self, on=True, skin_fraction=0.0):
        """Candidate lists in the contact kernel (plane-only batches); skin as a
        fraction of the cell size (0: keep the current one, 0.02 initially)."""
        abi.check(self.lib, self.lib.xdem_batch_set_lists(self.h, int(bool(on)),
                                                          float(skin_fraction)))

    def list_rebuilds(self):
        """Scenes whose candidate lists were all rebuilt for the last step."""
        m = C.c_uint32()
        abi.check(self.lib, self.lib.xdem_batch_list_rebuilds(self.h, C.byref(m)))
        return m.value

    def sort_patch(self, on=True):
        abi.check(self.lib, self.lib.xdem_batch_set_sort_patch(self.h, int(bool(on))))

    def autotune(self, steps_per_trial=20):
        """Pick the launch shape by timing whole graph-replayed step blocks
        (advances the scenes).  Returns (chosen code, {shape: ms per step})."""
        chosen = C.c_uint32()
        ms = np.zeros(54)
        abi.check(self.lib, self.lib.xdem_batch_autotune(self.h, steps_per_trial, C.byref(chosen),
                                                         abi.ptr(ms)))
        keys = [f"{w}-block{b}-unroll{u}{'-ahead' if p else ''}"
                for w in ("rows", "lane", "lane-lists")
                for b in (64, 128, 256) for u in (1, 2, 4) for p in (0, 1)]
        return chosen.value, dict(zip(keys, ms.tolist()))

    def state(self, sid):
        n = self.sizes[sid]
        out = dict(pos=np.empty((n, 4)), vel=np.empty((n, 3)), omega=np.empty((n, 3)),
                   acc=np.empty((n, 3)), alpha=np.empty((n, 3)))
        abi.check(self.lib, self.lib.xdem_batch_get_state(
            self.h, sid, *(abi.ptr(out[k]) for k in ("pos", "vel", "omega", "acc", "alpha"))))
        return out

    def history(self, sid):
        """(count[n], key[n, slots], xi[n, slots, 3]) in original particle order."""
        n = self.sizes[sid]
        count = np.zeros(n, np.uint32)
        key = np.zeros((n, HISTORY_SLOTS), np.uint32)
        xi = np.zeros((n, HISTORY_SLOTS, 3))
        abi.check(self.lib, self.lib.xdem_batch_get_history(self.h, sid, abi.ptr(count),
                                                            abi.ptr(key), abi.ptr(xi)))
        return count, key, xi

    def history_overflow(self):
        v = C.c_uint32()
        abi.check(self.lib, self.lib.xdem_batch_history_overflow(self.h, C.byref(v)))
        return v.value

    def energy(self, sid):
        ke = C.c_double()
        mom = np.zeros(3)
        abi.check(self.lib, self.lib.xdem_batch_energy(self.h, sid, C.byref(ke), abi.ptr(mom)))
        return ke.value, mom

    def contact_metrics(self, sid):
        """(particle-particle contact pairs, largest overlap) of a scene at its
        current positions; call prepare() first."""
        pairs = C.c_uint64()
        ov = C.c_double()
        abi.check(self.lib, self.lib.xdem_batch_contact_metrics(self.h, sid, C.byref(pairs),
                                                                C.byref(ov)))
        return pairs.value, ov.value

    def enable_reactions(self, on: bool = True):
        """Evaluate the mesh objects' reactions at the last step of every step call."""
        abi.check(self.lib, self.lib.xdem_batch_enable_reactions(self.h, int(on)))

    def _reactions(self, fn, sid):
        k = C.c_uint32()
        abi.check(self.lib, fn(self.h, sid, None, None, 0, C.byref(k)))
        f = np.zeros((k.value, 3))
        m = np.zeros((k.value, 3))
        if k.value:
            abi.check(self.lib, fn(self.h, sid, abi.ptr(f), abi.ptr(m), k.value, C.byref(k)))
        return f, m

    def object_reactions(self, sid):
        """(force[k, 3], moment[k, 3]) of the scene's k mesh objects at the last step:
        minus the summed triangle contact forces, and their moment about the
        object origin."""
        return self._reactions(self.lib.xdem_batch_object_reactions, sid)

    def drum_reactions(self, sid):
        """(force[k, 3], moment[k, 3]) of the scene's k drum walls at the last step,
        the moment about each drum's origin."""
        return self._reactions(self.lib.xdem_batch_drum_reactions, sid)

    def timing(self, on: bool):
        abi.check(self.lib, self.lib.xdem_batch_enable_timing(self.h, int(on)))

    def phase_times(self):
        ms = np.zeros(6)
        cnt = np.zeros(6, np.uint32)
        abi.check(self.lib, self.lib.xdem_batch_phase_times(self.h, abi.ptr(ms), abi.ptr(cnt)))
        return ms, cnt

    def last_movers(self):
        m = C.c_uint32()
        abi.check(self.lib, self.lib.xdem_batch_last_movers(self.h, C.byref(m)))
        return m.value

    def prepare(self):
        abi.check(self.lib, self.lib.xdem_batch_prepare(self.h))

    def capture(self, nsteps):
        """After prepare(): instantiate the graph of an nsteps block without
        running it (setup kept out of a timed run)."""
        abi.check(self.lib, self.lib.xdem_batch_capture(self.h, int(nsteps)))

    def probe(self, sid, ncell, max_per=32):
        n = self.sizes[sid]
        out = dict(force=np.zeros((n, 3)), moment=np.zeros((n, 3)),
                   counts=np.zeros(n, np.uint32), partners=np.zeros((n, max_per), np.uint32),
                   sorted_keys=np.zeros(n, np.uint32), sorted_ids=np.zeros(n, np.uint32),
                   cstart=np.zeros(ncell, np.uint32), cend=np.zeros(ncell, np.uint32))
        abi.check(self.lib, self.lib.xdem_batch_probe(
            self.h, sid, abi.ptr(out["force"]), abi.ptr(out["moment"]), max_per,
            abi.ptr(out["counts"]), abi.ptr(out["partners"]), abi.ptr(out["sorted_keys"]),
            abi.ptr(out["sorted_ids"]), abi.ptr(out["cstart"]), abi.ptr(out["cend"])))
        return out

    def close(self):
        if self.h:
            self.lib.xdem_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def apply_shape(b, chosen):
    """Set the launch shape an autotune code names (block size + 1000 * batch
    + 10000 * loaded ahead + 100000 * per-lane walk + 1000000 * lists)."""
    b.launch(chosen % 1000, (chosen // 1000) % 10, bool((chosen // 10000) % 10))
    b.walk(bool((chosen // 100000) % 10))
    b.lists(bool(chosen // 1000000))


def tune_shape(scenes_, lib, warmup=20, steps_per_trial=20, vels=None):
    """Autotune on a separate batch holding the same scenes (advanced `warmup`
    steps first), so the timed run starts from the workload's own state.
    Returns (chosen code, {shape: ms per step})."""
    n = sum(s.np for s in scenes_)
    cells = sum(int(np.prod(s.grid_size)) for s in scenes_)
    b = Batch(n, len(scenes_), cells, lib)
    for k, s in enumerate(scenes_):
        b.add(s, None if vels is None else vels[k])
    b.step(max(1, warmup))
    tuned = b.autotune(steps_per_trial=steps_per_trial)
    b.close()
    return tuned


def bench_fused(scene, steps, warmup, lib, block=100, graphs=True, autotune=True):
    """Steps `warmup`..`warmup + steps` of one scene on the fused path, timed
    (the scene resident in HBM first; graph-replayed blocks of `block` steps;
    the launch shape autotuned on a separate copy of the scene), then an eager
    pass with HIP-event phase timing for the roofline."""
    tuned = tune_shape([scene], lib) if autotune else None
    b = Batch(scene.np, 1, int(np.prod(scene.grid_size)), lib)
    b.graphs(graphs)
    if tuned:
        apply_shape(b, tuned[0])
    sid = b.add(scene)
    blk = min(block, max(1, steps))
    b.prepare()  # the scene resident in HBM and sorted before the timed steps
    b.capture(blk)  # the block's graph instantiated before the timed steps
    if steps % blk:
        b.capture(steps % blk)
    done = 0
    while done < warmup:
        k = min(blk, warmup - done)
        b.step(k)
        done += k
    b.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(blk, steps - done)
        b.step(k)
        done += k
    b.sync()
    wall = time.perf_counter() - t0
    # eager pass with phase timing (same kernels, events around each phase)
    b.timing(True)
    nt = min(steps, 50)
    b.step(nt)
    ms, cnt = b.phase_times()
    b.timing(False)
    st = b.state(sid)
    assert np.isfinite(st["pos"]).all()
    movers = b.last_movers()
    b.close()
    per = dict(zip(PHASES, (ms / nt).tolist()))
    # the same job from host buffers: scene upload (host -> HBM), the steps
    # (first block captured), state read back (HBM -> host)
    b = Batch(scene.np, 1, int(np.prod(scene.grid_size)), lib)
    b.graphs(graphs)
    if tuned:
        apply_shape(b, tuned[0])
    b.sync()
    t0 = time.perf_counter()
    sid = b.add(scene)
    done = 0
    while done < steps:
        k = min(blk, steps - done)
        b.step(k)
        done += k
    st = b.state(sid)
    pcie_wall = time.perf_counter() - t0
    b.close()
    return dict(wall_s=wall, ms_per_step=wall * 1e3 / steps, value=scene.np * steps / wall,
                phase_ms=per, movers_last=movers, timed_steps_events=nt,
                autotune=None if tuned is None else {"chosen": tuned[0], "step_ms": tuned[1]},
                pcie_inclusive={"value": scene.np * steps / pcie_wall, "wall_s": pcie_wall,
                                "steps": steps})


def smoke_check(scene, lib, w_pos=None):
    b = Batch(scene.np, 1, int(np.prod(scene.grid_size)), lib)
    sid = b.add(scene)
    b.step(20)
    st = b.state(sid)
    assert np.isfinite(st["pos"]).all()
    if w_pos is not None:
        err = np.abs(st["pos"] - w_pos).max()
        print(f"[smoke] fused FP32c vs oracle after 20 steps: max |dpos| = {err:.3e} m", flush=True)
        assert err < 1e-6, err
    b.close()
