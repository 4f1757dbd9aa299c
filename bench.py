"""Benchmark of the DEM contact step on one MI355X.

Workload (default): the 1M-sphere polydisperse hopper (radii U[2,4] mm, 6
planes, dt = 5e-6 s, cell 8 mm, 128^3 grid), synthetic seeded fill.  One
"step" = one full DEM time step (Verlet position, hash, sort, cell ranges,
particle-particle and wall contacts, Verlet velocity) over all particles.

Prints ONE JSON line (rank 0, N = 1).  `value` is the throughput of the fused
FP32c path when it is built, else of the plain FP64 path; both are reported.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from metakonga_amd import abi, scenes  # noqa: E402

METRIC = "particle-steps/s (FP32-comp batched vs plain FP64) + p50/p99 scene latency"
HBM_PEAK = 8.0e12
PHASES = ["vv_position", "hash", "sort", "cell_ranges", "p2p", "walls", "triangles", "vv_velocity"]
# Algorithmic bytes per particle of the plain FP64 particle-particle kernel:
# own pos+r 32, vel 24, omega 24, mass 8, sorted index 4, force and moment
# read+write 96 (neighbour re-reads excluded).
P2P_F64_BYTES = 188.0
# Algorithmic bytes per particle-step of the whole FP64 step (reference data model).
STEP_F64_BYTES = 300.0
# Algorithmic bytes per particle of the fused contact + Verlet kernel (k_step):
# own state record read 88 (pos+r 32, v+m 16, omega+alpha.x 16, a+alpha.y 16,
# alpha.z+id 8) + written 88 + FP32 position copy written 16 + sorted->storage
# index 4 + next cell key written 4 + previous cell key read 4 (mover
# detection for the incremental sort); neighbour re-reads excluded.
KSTEP_BYTES = 204.0
# Algorithmic bytes per particle-step of the whole FP32c step.
STEP_F32C_BYTES = 192.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_hopper1M_fp32c.md")


def profiled_traffic(kernel_sig):
    """HBM bytes per launch of `kernel_sig` (e.g. "k_step<true, false, 2, false>")
    from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (counter values
    scaled by the copy-kernel calibration); None when the table has no row."""
    try:
        for line in open(TRAFFIC_TABLE):
            cols = [c.strip() for c in line.split("|")]
            if len(cols) > 7 and cols[1].strip("`").endswith(kernel_sig):
                return (float(cols[5]) + float(cols[6])) * 1e6
    except (OSError, ValueError):
        pass
    return None


def run_fp64(scene, steps, warmup, lib, phase_steps=20):
    """The plain FP64 path: `steps` steps after `warmup` (timed without
    events), then `phase_steps` more with HIP-event phase timing."""
    from metakonga_amd.xdem import Simulation
    sim = Simulation(scene, lib)
    t0 = time.perf_counter()
    if warmup:
        sim.step(warmup)
    sim.sync()
    log(f"fp64 warmup {warmup} steps {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    abi.check(lib, lib.xdem_sim_enable_timing(sim.h, 1))
    sim.step(phase_steps)
    sim.sync()
    ms = np.zeros(8)
    cnt = np.zeros(8, np.uint32)
    abi.check(lib, lib.xdem_sim_phase_times(sim.h, abi.ptr(ms), abi.ptr(cnt)))
    abi.check(lib, lib.xdem_sim_enable_timing(sim.h, 0))
    # untimed check: the run stayed finite
    st = sim.state()
    assert np.isfinite(st["pos"]).all()
    sim.close()
    n = scene.np
    return dict(wall_s=wall, ms_per_step=wall * 1e3 / steps, value=n * steps / wall,
                phase_ms=dict(zip(PHASES, (ms / phase_steps).tolist())),
                phase_launches=dict(zip(PHASES, cnt.tolist())), phase_steps=phase_steps)


def run_batched(nscene, steps, lib, block=100, seed=0xD3E3):
    """Batched box settling (many independent scenes, N_s ~ logU[2k, 200k],
    one radius per scene), all advanced together on the fused path: the whole
    `steps`-step job from the scenes' initial state (resident in HBM first),
    the launch shape autotuned on a separate batch of the same scenes."""
    from metakonga_amd import batch
    t0 = time.perf_counter()
    scs = scenes.batch_scenes(nscene=nscene, seed=seed, nsteps=steps)
    gen = time.perf_counter() - t0
    ntot = sum(s.np for s in scs)
    cells = sum(int(np.prod(s.grid_size)) for s in scs)
    tuned = batch.tune_shape(scs, lib)
    b = batch.Batch(ntot, nscene, cells, lib)
    batch.apply_shape(b, tuned[0])
    for s in scs:
        b.add(s)
    b.prepare()
    b.capture(block)
    b.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        b.step(block)
        done += block
    b.sync()
    wall = time.perf_counter() - t0
    b.close()
    return dict(scenes=nscene, particles=ntot, steps=done, cells=cells, wall_s=wall,
                value=ntot * done / wall, generation_s=gen,
                autotune={"chosen": tuned[0], "step_ms": tuned[1]})


def run_settling_8k(lib, steps=2000):
    """The 8,000-sphere settling box (r = 5 mm, dt = 1e-5 s, 2,000 steps) on both
    paths: small enough that launch latency, not bandwidth, sets the rate."""
    from metakonga_amd import batch
    from metakonga_amd.xdem import Simulation
    sc = scenes.settling_box()
    sim = Simulation(sc, lib)
    sim.step(10)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    f64 = sc.np * steps / (time.perf_counter() - t0)
    sim.close()
    tuned = batch.tune_shape([sc], lib)
    b = batch.Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
    batch.apply_shape(b, tuned[0])
    b.add(sc)
    b.prepare()
    b.capture(100)
    b.sync()
    t0 = time.perf_counter()
    for _ in range(steps // 100):
        b.step(100)
    b.sync()
    f32c = sc.np * steps / (time.perf_counter() - t0)
    b.close()
    return {"np": sc.np, "steps": steps, "fp64_plain": f64, "fp32c_fused": f32c}


def run_c2_drum(lib, steps=200, f64_steps=20):
    """The drum workload: 4e6 spheres (r = 3 mm) in the 30 rpm drum with its turning
    20,000-triangle lifter-plate baffle, dt = 5e-6 s, on the fused FP32c path (graph-replayed
    100-step blocks, autotuned) and on the plain FP64 path.  Informational: a
    reduced number of steps of the 2,000-step configuration."""
    from metakonga_amd import batch
    from metakonga_amd.xdem import Simulation
    t0 = time.perf_counter()
    sc = scenes.drum()
    gen = time.perf_counter() - t0
    tuned = batch.tune_shape([sc], lib)
    b = batch.Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
    batch.apply_shape(b, tuned[0])
    sid = b.add(sc)
    b.prepare()
    b.capture(100)
    b.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(100, steps - done)
        b.step(k)
        done += k
    b.sync()
    f32c = sc.np * steps / (time.perf_counter() - t0)
    b.timing(True)
    b.step(20)
    ms, cnt = b.phase_times()
    b.timing(False)
    st = b.state(sid)
    assert np.isfinite(st["pos"]).all()
    b.close()
    sim = Simulation(sc, lib)
    sim.step(3)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(f64_steps)
    sim.sync()
    f64 = sc.np * f64_steps / (time.perf_counter() - t0)
    sim.close()
    return {"workload": f"drum {sc.np} spheres r 3 mm, R 0.5 m, L 1.6 m, 30 rpm, "
                        f"{len(sc.tri)}-triangle turning lifter plate, dt 5e-6 s, cell 6 mm, grid 256^3",
            "fp32c_fused": f32c, "fp32c_steps": steps, "fp64_plain": f64, "fp64_steps": f64_steps,
            "speedup_vs_fp64": f32c / f64, "contact_ms": ms[1] / cnt[1],
            "autotune_chosen": tuned[0], "generation_s": gen}


def run_serving(duration, lib, rate=20.0, block=100, seed=0xD3E4):
    """Continuous batching of a seeded Poisson trace (drop tests and hopper
    discharges, 1k-500k spheres, 500-5000 steps); scenes are generated before
    the trace starts (a request carries its scene)."""
    from metakonga_amd import batch, serve
    trace = serve.poisson_trace(seed=seed, rate=rate, duration=duration, block=block)
    cache = {}
    for r in trace:
        cache[r.rid] = r.scene()
    for r in trace:
        r.scene = (lambda rr: (lambda: cache[rr.rid]))(r)
    cap_np = 3_000_000
    b = batch.Batch(cap_np, 128, 1 << 28, lib)
    srv = serve.Server(b, max_particles=cap_np, max_scenes=128, max_cells=1 << 28,
                       hbm_budget=64 << 30, block=block)
    t0 = time.perf_counter()
    srv.run(trace, t0=t0)
    wall = time.perf_counter() - t0
    rep = srv.report(wall)
    b.close()
    rep.update(trace_s=duration, rate_per_s=rate, wall_s=wall,
               offered_particle_steps_per_s=sum(r.n * r.steps for r in trace) / duration)
    return rep


def fp32c_error_vs_oracle(lib, n=1_000_000, steps=300, speed=1.0, sc=None, label=None):
    """Relative force and torque error of the fused FP32c contact evaluation
    against the FP64 oracle on one state of the hopper workload (or of `sc`):
    the seeded fill given random velocities of `speed` m/s (none for `sc`) and
    advanced `steps` steps on the fused path (so that spheres collide), then
    every particle's contact force and torque evaluated by both on that state.
    Per particle the error is |F_gpu - F_oracle|_inf / (sum_k |F_k| + m g), the
    sum over its contacts as the oracle evaluates them; the maximum over
    particles is reported."""
    from metakonga_amd import batch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.load()
    if sc is None:
        sc = scenes.hopper(n=n)
        vel = scenes.as_f32_exact(scenes.uniform(0xABC, 3 * n, 5).reshape(n, 3) * 2.0 * speed - speed)
        label = (f"hopper {n} spheres after {steps} steps from the fill with random "
                 f"velocities of {speed} m/s (FP32c path), forces of every particle")
    else:
        n = sc.np
        vel = None
    ncell = int(np.prod(sc.grid_size))
    b = batch.Batch(n, 1, ncell, lib)
    sid = b.add(sc, vel)
    done = 0
    while done < steps:
        b.step(min(100, steps - done))
        done += 100
    st = b.state(sid)
    b.prepare()
    out = b.probe(sid, ncell, max_per=32)
    b.close()
    p = sc.params()
    pos = np.ascontiguousarray(st["pos"])
    vv, om = np.ascontiguousarray(st["vel"]), np.ascontiguousarray(st["omega"])
    h, idx = pyoracle.hash_and_index(p, pos)
    h, idx, cs, ce, srt = pyoracle.reorder(h, idx, p.ncell)
    m, _ = sc.mass_inertia()
    f, t = np.zeros((n, 3)), np.zeros((n, 3))
    with pyoracle.force_magnitudes(n) as fm:
        npairs = pyoracle.p2p(p, 1, pos, vv, om, f, t, m, srt, cs, ce, sc.contact_pp())
        for pl in sc.planes:
            pyoracle.plane(p, 1, pl, pos, vv, om, f, t, m, sc.contact_pp())
    wg = m * 9.80665
    df = np.abs(out["force"] - f).max(1)
    dt = np.abs(out["moment"] - t).max(1)
    st, sn = fm.terms + wg, np.asarray(fm) + wg
    en = df / sn
    exact = bool(np.array_equal(out["sorted_ids"], srt) and np.array_equal(out["cstart"], cs) and
                 np.array_equal(out["cend"], ce) and int(out["counts"].sum()) == int(npairs))
    return {"force_rel_err_max": float((df / st).max()),
            "torque_rel_err_max": float((dt / (st * pos[:, 3])).max()),
            "scale": "per particle: sum over its contacts of |fsn| + |fca| + (|vn| + |vs|) V + "
                     "|Ft| (the contact-law term magnitudes) + m g",
            "force_rel_err_max_net_scale": float(en.max()),
            "force_rel_err_p9999_net_scale": float(np.quantile(en, 0.9999)),
            "net_scale": "per particle: sum over its contacts of |Fn + Ft| + m g",
            "share_within_1e-5_net_scale": float((en <= 1e-5).mean()),
            "contact_pairs": int(npairs), "contacts_per_particle": int(npairs) / n,
            "grid_and_pair_count_exact": exact, "state": label}


def run_dense_bed(lib, steps=200, f64_steps=20, err=True):
    """Contact-rich leg: 1M spheres (r U[3.92, 4] mm) in a compressed
    face-centred cubic packing in a closed box (about 9-10 contacts per sphere,
    solid fraction about 0.7), on the fused FP32c path (autotuned, graph
    blocks; kernel time from HIP events for the roofline) and on the plain
    FP64 path, with the FP32c force error against the FP64 oracle on the state
    after 100 steps."""
    from metakonga_amd import batch
    sc = scenes.dense_bed()
    fused = batch.bench_fused(sc, steps, 20, lib)
    k_ms = fused["phase_ms"]["contact_verlet"]
    f64 = run_fp64(sc, f64_steps, 3, lib)
    out = {"workload": f"dense bed {sc.np} spheres, FCC lattice at 0.985 x 2r, radii "
                       f"U[3.92, 4] mm, closed 6-plane box, dt 5e-6 s, cell 8 mm, "
                       f"grid {sc.grid_size[0]}^3",
           "fp32c_fused": fused["value"], "ms_per_step": fused["ms_per_step"],
           "phase_ms": fused["phase_ms"], "autotune_chosen": (fused.get("autotune") or {}).get("chosen"),
           "fp64_plain": f64["value"], "fp64_steps": f64_steps,
           "speedup_vs_fp64": fused["value"] / f64["value"],
           "roofline": {"bound": "hbm", "kernel_ms": k_ms,
                        "achieved": STEP_F32C_BYTES * sc.np / (k_ms * 1e-3) / 1e9,
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": STEP_F32C_BYTES * sc.np / (k_ms * 1e-3) / HBM_PEAK,
                        "step_frac": STEP_F32C_BYTES * fused["value"] / HBM_PEAK}}
    if err:
        out["fp32c_force_err_vs_oracle"] = fp32c_error_vs_oracle(
            lib, steps=100, sc=sc, label=f"dense bed {sc.np} spheres after 100 steps (FP32c "
                                         f"path), forces of every particle")
    return out


def cpu_subset(scs, budget_s, threads=1):
    """Oracle CPU path over a fixed subset of scenes: each scene stepped for
    its share of the time budget; particle-steps/s over the whole subset."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.load()
    total_ps, total_t, steps = 0, 0.0, []
    for sc in scs:
        w = pyoracle.World(sc, semantics=pyoracle.ORC_HOST)
        t0 = time.perf_counter()
        w.step(1, threads=threads)
        t1 = time.perf_counter() - t0
        k = max(1, min(200, int(budget_s / len(scs) / max(t1, 1e-4)) - 1))
        t0 = time.perf_counter()
        w.step(k, threads=threads)
        total_t += time.perf_counter() - t0
        total_ps += sc.np * k
        steps.append(k)
    return dict(value=total_ps / total_t, seconds=total_t, particle_steps=total_ps,
                steps_min=min(steps), steps_max=max(steps))


def cpu_baseline(scene, budget_s=15.0, threads=1):
    """Oracle CPU path (serial, the reference's CPU semantics) on a bounded
    sample of the same workload: all particles, as many steps as fit ~budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.load()
    w = pyoracle.World(scene, semantics=pyoracle.ORC_HOST)
    t0 = time.perf_counter()
    w.step(1, threads=threads)
    t1 = time.perf_counter() - t0
    k = max(1, min(50, int(budget_s / max(t1, 1e-3)) - 1))
    t0 = time.perf_counter()
    w.step(k, threads=threads)
    dt = time.perf_counter() - t0
    return dict(value=scene.np * k / dt, steps=k, seconds=dt)


def cpu_settling_full():
    """The 8,000-sphere settling box in full (2,000 steps, 1.6e7 particle-steps)
    on the oracle's serial CPU-path semantics."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.load()
    sc = scenes.settling_box()
    w = pyoracle.World(sc, semantics=pyoracle.ORC_HOST)
    t0 = time.perf_counter()
    w.step(sc.nsteps)
    dt = time.perf_counter() - t0
    return dict(value=sc.np * sc.nsteps / dt, steps=sc.nsteps, seconds=dt, np=sc.np)


def cpu_model():
    """lscpu's full model name, sockets and threads, and nproc."""
    info = {"nproc": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout)
    except Exception:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Thread(s) per core", "Core(s) per socket",
                             "CPU(s)"):
                info[k.strip()] = v.strip()
    except Exception:
        info["Model name"] = platform.processor()
    return info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000, help="timed steps (the whole 5,000-step run)")
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--np", type=int, default=1_000_000, help="particles of the hopper workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--batched-scenes", type=int, default=256, help="0 to skip the batched run")
    ap.add_argument("--batched-steps", type=int, default=1000)
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="0 to skip serving")
    ap.add_argument("--skip-fp64", action="store_true", help="skip the plain FP64 leg")
    ap.add_argument("--dense-steps", type=int, default=200,
                    help="steps of the contact-rich 1M dense-bed leg (0 to skip)")
    ap.add_argument("--c2-steps", type=int, default=200,
                    help="steps of the 4e6-sphere drum leg (0 to skip)")
    ap.add_argument("--no-graphs", action="store_true", help="eager launches (profiler runs)")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shape")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run the plain copy kernel on this many MiB, 20 times (profiler "
                         "byte-counter calibration), then exit")
    ap.add_argument("--block-log", default="",
                    help="only run the hopper with per-block JSON metrics (contacts per particle, "
                         "largest overlap, kinetic energy, momentum) written to this file, then exit")
    ap.add_argument("--block-log-steps", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    lib = abi.load()
    if lib.xdem_device_count() < 1:
        raise SystemExit("no GPU visible to libxdem.so")
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src, dst = C.c_void_p(), C.c_void_p()
        abi.check(lib, lib.xdem_malloc(C.byref(src), nb))
        abi.check(lib, lib.xdem_malloc(C.byref(dst), nb))
        abi.check(lib, lib.xdem_memset(src, 1, nb))
        for _ in range(20):
            abi.check(lib, lib.xdem_copy_kernel(dst, src, nb))
        abi.check(lib, lib.xdem_synchronize())
        print(json.dumps({"copy_calibration_bytes": nb, "launches": 20}))
        return
    if args.block_log:
        from metakonga_amd import batch, blocklog
        sc = scenes.hopper(n=args.np)
        b = batch.Batch(sc.np, 1, int(np.prod(sc.grid_size)), lib)
        b.lists(True)
        sid = b.add(sc)
        b.prepare()
        with open(args.block_log, "w") as f:
            recs = blocklog.run_logged(b, [sid], args.block_log_steps, 100, f)
        b.close()
        print(json.dumps({"block_log": args.block_log, "blocks": len(recs), "last": recs[-1]}))
        return
    t0 = time.perf_counter()
    scene = scenes.hopper(n=args.np)
    log(f"scene {scene.name} np={scene.np} grid={scene.grid_size} cs={scene.cell_size:.4g} "
        f"built in {time.perf_counter() - t0:.1f}s")

    f64 = None
    if not args.skip_fp64:
        f64 = run_fp64(scene, args.steps, args.warmup, lib)
        log(f"fp64: {f64['value']:.4g} particle-steps/s, {f64['ms_per_step']:.3f} ms/step")
        log("fp64 phases ms/step: " + ", ".join(f"{k}={v:.4f}" for k, v in f64["phase_ms"].items()))

    fused = None
    try:
        from metakonga_amd import batch
        if batch.available(lib):
            fused = batch.bench_fused(scene, args.steps, args.warmup, lib,
                                      graphs=not args.no_graphs, autotune=not args.no_autotune)
            log(f"fp32c: {fused['value']:.4g} particle-steps/s, {fused['ms_per_step']:.3f} ms/step")
    except ImportError:
        fused = None

    if fused is not None:
        head = fused
        dtype = "f32 contact + f64 positions"
        k_ms = fused["phase_ms"]["contact_verlet"]
        # algorithmic bytes of one launch: the per-particle-step figure of the
        # FP32c data model (192 B) x the particles one launch advances
        achieved = STEP_F32C_BYTES * scene.np / (k_ms * 1e-3)
        ch = (fused.get("autotune") or {}).get("chosen", 2256)
        tf = lambda v: "true" if v else "false"  # noqa: E731
        sig = (f"k_step<true, false, {(ch // 1000) % 10}, {tf((ch // 10000) % 10)}, false, false, "
               f"{tf((ch // 100000) % 10)}, {tf(ch // 1000000)}>")
        traffic = profiled_traffic(sig) if scene.np == 1_000_000 else None
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic,
                "traffic_source": "profiles/hbm_traffic_hopper1M_fp32c.md (rocprofv3 FETCH_SIZE + "
                                  "WRITE_SIZE, copy-kernel calibrated)" if traffic else None,
                "kernel": f"{sig} (fused contact + Verlet, FP32c), block {ch % 1000}",
                "bytes_per_particle": STEP_F32C_BYTES, "kernel_record_bytes": KSTEP_BYTES,
                "kernel_ms": k_ms,
                "step_frac": STEP_F32C_BYTES * fused["value"] / HBM_PEAK}
    else:
        head = f64
        dtype = "f64"
        p2p_ms = f64["phase_ms"]["p2p"]
        achieved = P2P_F64_BYTES * scene.np / (p2p_ms * 1e-3)
        dom_kernel = "k_p2p<1> (plain FP64 particle-particle)"
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None,
                "kernel": dom_kernel, "bytes_per_particle": P2P_F64_BYTES,
                "kernel_ms": p2p_ms}

    batched = None
    if fused is not None and args.batched_scenes > 0:
        batched = run_batched(args.batched_scenes, args.batched_steps, lib)
        log(f"batched: {batched['scenes']} scenes {batched['particles']} particles: "
            f"{batched['value']:.4g} particle-steps/s")
    serving = None
    if fused is not None and args.serve_seconds > 0:
        serving = run_serving(args.serve_seconds, lib)
        log(f"serving: {serving}")

    small = None
    if fused is not None and not args.skip_fp64 and args.serve_seconds > 0:
        small = run_settling_8k(lib)
        log(f"settling 8k: {small}")

    dense = None
    if fused is not None and not args.skip_fp64 and args.dense_steps > 0:
        dense = run_dense_bed(lib, args.dense_steps, err=not args.no_cpu)
        log(f"dense bed: {dense}")

    c2 = None
    if fused is not None and not args.skip_fp64 and args.c2_steps > 0:
        c2 = run_c2_drum(lib, args.c2_steps)
        log(f"C2 drum: {c2}")

    cpu = None
    cpu_omp = None
    cpu_small = None
    cpu_c3 = cpu_c4 = None
    err = None
    if not args.no_cpu:
        host = cpu_model()
        log(f"host: {host}")
        if fused is not None and args.np == 1_000_000:
            err = fp32c_error_vs_oracle(lib)
            log(f"fp32c vs fp64 oracle: {err}")
        log("cpu baseline (oracle, CPU-path semantics, 1 thread) ...")
        c = cpu_baseline(scene, budget_s=args.cpu_budget)
        cpu = {"value": c["value"], "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"{scene.name} all {scene.np} particles, {c['steps']} steps "
                         f"({c['seconds']:.1f}s), serial CPU-path semantics",
               "cpu": host}
        log(f"cpu: {c['value']:.4g} particle-steps/s over {c['steps']} steps")
        nthr = host.get("nproc") or os.cpu_count() or 1
        co = cpu_baseline(scene, budget_s=args.cpu_budget / 2, threads=nthr)
        cpu_omp = {"value": co["value"], "unit": "particle-steps/s", "cores": nthr,
                   "kind": "port", "sample": f"same, OpenMP over particles on all {nthr} host "
                                             f"threads, {co['steps']} steps"}
        log(f"cpu omp x{nthr}: {co['value']:.4g} particle-steps/s")
        cs = cpu_settling_full()
        cpu_small = {"value": cs["value"], "unit": "particle-steps/s", "cores": 1, "kind": "port",
                     "sample": f"settling box {cs['np']} spheres in full, {cs['steps']} steps "
                               f"({cs['seconds']:.1f}s), serial CPU-path semantics"}
        log(f"cpu settling 8k in full: {cs['value']:.4g} particle-steps/s")
        from metakonga_amd import serve
        c3 = cpu_subset(scenes.batch_scenes()[:16], args.cpu_budget / 2, threads=nthr)
        cpu_c3 = {"value": c3["value"], "unit": "particle-steps/s", "cores": nthr, "kind": "port",
                  "sample": f"the first 16 of the 256 box-settling scenes, "
                            f"{c3['steps_min']}-{c3['steps_max']} steps each "
                            f"({c3['seconds']:.1f}s), OpenMP over particles"}
        tr = serve.poisson_trace(seed=0xD3E4, rate=20.0, duration=60.0, block=100)[:16]
        c4 = cpu_subset([r.scene() for r in tr], args.cpu_budget / 2, threads=nthr)
        cpu_c4 = {"value": c4["value"], "unit": "particle-steps/s", "cores": nthr, "kind": "port",
                  "sample": f"the first 16 requests of the serving trace, "
                            f"{c4['steps_min']}-{c4['steps_max']} steps each "
                            f"({c4['seconds']:.1f}s), OpenMP over particles"}
        log(f"cpu 16-scene subsets: batched {c3['value']:.4g}, serving {c4['value']:.4g}")

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded jittered-lattice fill, radii U[2,4] mm)",
        "config": {"workload": f"hopper {scene.np} polydisperse spheres, 6 planes, dt 5e-6 s, "
                               f"cell 8 mm, grid {scene.grid_size[0]}^3",
                   "np": scene.np, "path": "fp32c-fused" if fused else "fp64-plain"},
        "fp64_plain": None if f64 is None else {
            "value": f64["value"], "ms_per_step": f64["ms_per_step"], "phase_ms": f64["phase_ms"],
            "step_roofline_frac": STEP_F64_BYTES * f64["value"] / HBM_PEAK},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if fused is not None:
        out["fp32c_fused"] = {k: v for k, v in fused.items() if k not in ("roofline",)}
        if f64 is not None:
            out["speedup_vs_fp64"] = fused["value"] / f64["value"]
    if batched is not None:
        out["batched"] = batched
    if small is not None:
        out["settling_8k"] = small
    if dense is not None:
        out["dense_bed"] = dense
    if c2 is not None:
        out["c2_drum"] = c2
    if cpu_omp is not None:
        out["cpu_baseline_multicore"] = cpu_omp
    if cpu_small is not None:
        out["cpu_baseline_settling_8k"] = cpu_small
    if cpu_c3 is not None:
        out["cpu_baseline_batched_16"] = cpu_c3
        out["cpu_baseline_serving_16"] = cpu_c4
    if err is not None:
        out["fp32c_force_err_vs_oracle"] = err
    if serving is not None:
        out["serving"] = serving
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
