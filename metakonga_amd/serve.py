"""Continuous-batching request loop for DEM scenes on one GPU.

Requests (drop-test boxes and hopper discharges of different sizes and step
counts) arrive over time.  The loop keeps every admitted scene's state
resident in one fused FP32c batch (metakonga_amd.batch.Batch), advances all of
them together in step blocks, admits waiting requests at block boundaries as
soon as particle slots, cell-key space and the HBM budget allow, and retires a
scene when its steps are done (its result is read back, its slot freed).

The backend is any object with the Batch interface (add/remove/step/sync/
state); tests drive the host logic with a fake backend and a fake clock.
No counterpart in the reference (its driver runs one scene,
source/xDynamicsSolver.cpp:132-252).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, scenes


def device_bytes(particles, nscenes, cells, history=False, lists=True, lib=None):
    """Device bytes of a fused batch holding `particles`, `nscenes` scenes and
    `cells` cells (with candidate lists unless lists=False): the library's own
    allocation arithmetic (xdem_batch_device_bytes; host only, no GPU needed)."""
    from . import batch as _batch
    L = lib or abi.load()
    _batch.declare(L)
    return int(L.xdem_batch_device_bytes(int(particles), int(nscenes), int(cells),
                                         (1 if history else 0) | (2 if lists else 0)))


@dataclass
class Request:
    rid: int
    arrival: float           # seconds since the start of the trace
    kind: str                # "drop" or "hopper"
    n: int                   # particles
    steps: int               # multiple of the block size
    seed: int

    def scene(self):
        if self.kind == "hopper":
            return scenes.hopper(n=self.n, seed=self.seed, nsteps=self.steps, name=f"req{self.rid}")
        r = float(scenes.as_f32_exact(2e-3 + 3e-3 * scenes.uniform(self.seed, 1, 9)[0]))
        return scenes.box_scene(self.n, r, self.seed, 5e-6, self.steps, name=f"req{self.rid}")


@dataclass
class Record:
    req: Request
    admitted: float = math.nan
    finished: float = math.nan
    sid: int = -1
    done_steps: int = 0
    result: dict | None = None

    @property
    def latency(self):
        return self.finished - self.req.arrival


def poisson_trace(seed=0xD3E4, rate=20.0, duration=60.0, n_lo=1000, n_hi=500_000,
                  steps_lo=500, steps_hi=5000, block=100, max_requests=None):
    """Seeded Poisson arrivals (exponential gaps), half drop tests and half
    hopper discharges, N ~ logU[n_lo, n_hi], steps ~ U{steps_lo..steps_hi}
    rounded up to the block size."""
    out = []
    t = 0.0
    k = 0
    while True:
        u = scenes.uniform(seed, 5, stream=1000 + k)
        t += -math.log(1.0 - u[0]) / rate
        if t > duration or (max_requests and k >= max_requests):
            break
        n = int(math.exp(math.log(n_lo) + (math.log(n_hi) - math.log(n_lo)) * u[1]))
        steps = steps_lo + int(u[2] * (steps_hi - steps_lo + 1))
        steps = int(math.ceil(steps / block) * block)
        kind = "drop" if u[3] < 0.5 else "hopper"
        out.append(Request(k, t, kind, n, steps, seed * 7919 + k))
        k += 1
    return out


@dataclass
class Server:
    backend: object
    max_particles: int
    max_scenes: int
    max_cells: int
    hbm_budget: int          # bytes usable for scene state
    block: int = 100
    clock: object = time.perf_counter
    keep_results: bool = False
    waiting: list = field(default_factory=list)
    running: dict = field(default_factory=dict)
    finished: list = field(default_factory=list)
    used_particles: int = 0
    used_cells: int = 0
    used_history: bool = False
    particle_steps: int = 0

    def footprint(self, sc):
        """Device bytes of a batch holding this scene alone."""
        return device_bytes(sc.np, 1, int(np.prod(sc.grid_size)), sc.tangential_history)

    def fits(self, sc):
        cells = int(np.prod(sc.grid_size))
        mem = device_bytes(self.used_particles + sc.np, len(self.running) + 1,
                           self.used_cells + cells, self.used_history or sc.tangential_history)
        return (self.used_particles + sc.np <= self.max_particles and
                self.used_cells + cells <= self.max_cells and
                len(self.running) < self.max_scenes and mem <= self.hbm_budget)

    def admit(self, now):
        """First-come first-served: admit in arrival order while the head fits
        (a request that can never fit is rejected)."""
        while self.waiting:
            rec = self.waiting[0]
            sc = rec.req.scene()
            if (sc.np > self.max_particles or int(np.prod(sc.grid_size)) > self.max_cells or
                    self.footprint(sc) > self.hbm_budget):
                self.waiting.pop(0)
                rec.finished = now
                rec.result = {"rejected": True}
                self.finished.append(rec)
                continue
            if not self.fits(sc):
                break
            self.waiting.pop(0)
            rec.sid = self.backend.add(sc)
            rec.admitted = now
            rec.scene_np = sc.np
            rec.scene_cells = int(np.prod(sc.grid_size))
            self.used_particles += sc.np
            self.used_cells += rec.scene_cells
            self.used_history = self.used_history or sc.tangential_history
            self.running[rec.sid] = rec

    def retire(self, now):
        done = [r for r in self.running.values() if r.done_steps >= r.req.steps]
        for r in done:
            if self.keep_results:
                r.result = self.backend.state(r.sid)
            self.backend.remove(r.sid)
            r.finished = now
            self.used_particles -= r.scene_np
            self.used_cells -= r.scene_cells
            del self.running[r.sid]
            self.finished.append(r)

    def run(self, trace, t0=None):
        """Serve the trace; arrivals are mapped onto the clock (t0 + arrival)."""
        pending = sorted(trace, key=lambda r: r.arrival)
        t0 = self.clock() if t0 is None else t0
        i = 0
        while i < len(pending) or self.waiting or self.running:
            now = self.clock() - t0
            while i < len(pending) and pending[i].arrival <= now:
                self.waiting.append(Record(pending[i]))
                i += 1
            self.admit(now)
            if not self.running:
                if i < len(pending):   # idle until the next arrival
                    nxt = pending[i].arrival
                    self.idle_until(t0 + nxt)
                continue
            self.backend.step(self.block)
            self.backend.sync()
            for r in self.running.values():
                r.done_steps += self.block
                self.particle_steps += self.block * r.scene_np
            self.retire(self.clock() - t0)
        return self.finished

    def idle_until(self, t):
        dt = t - self.clock()
        if dt > 0 and self.clock is time.perf_counter:
            time.sleep(dt)
        elif dt > 0 and hasattr(self.clock, "advance"):
            self.clock.advance(dt)

    def report(self, wall):
        lat = np.array([r.latency for r in self.finished if not (r.result or {}).get("rejected")])
        return {"requests": len(self.finished),
                "rejected": sum(1 for r in self.finished if (r.result or {}).get("rejected")),
                "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
                "particle_steps": int(self.particle_steps),
                "particle_steps_per_s": self.particle_steps / wall if wall > 0 else None}


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def advance(self, dt):
        self.t += dt
