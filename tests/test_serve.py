"""Host logic of the continuous-batching request loop (CPU, fake backend and
fake clock): admission by capacity and HBM budget, FCFS, retirement, latency."""
import numpy as np

from metakonga_amd import serve


class FakeBackend:
    """Stands in for the GPU batch: a block of k steps costs k * particles * c seconds."""

    def __init__(self, clock, cost=2e-10):
        self.clock = clock
        self.cost = cost
        self.live = {}
        self.next = 0
        self.max_live = 0
        self.max_particles = 0

    def add(self, sc):
        sid = self.next
        self.next += 1
        self.live[sid] = sc.np
        self.max_live = max(self.max_live, len(self.live))
        self.max_particles = max(self.max_particles, sum(self.live.values()))
        return sid

    def remove(self, sid):
        del self.live[sid]

    def step(self, k):
        self.clock.advance(k * sum(self.live.values()) * self.cost)

    def sync(self):
        pass

    def state(self, sid):
        return {"np": self.live[sid]}


def _server(clock, backend, max_particles=600_000, budget=None):
    return serve.Server(backend, max_particles=max_particles, max_scenes=64, max_cells=1 << 28,
                        hbm_budget=budget or (1 << 40), block=100, clock=clock)


def test_trace_is_seeded_and_shaped():
    a = serve.poisson_trace(seed=5, rate=20, duration=10)
    b = serve.poisson_trace(seed=5, rate=20, duration=10)
    assert [(r.arrival, r.n, r.steps, r.kind) for r in a] == [(r.arrival, r.n, r.steps, r.kind) for r in b]
    assert 120 < len(a) < 280                     # ~ rate * duration
    assert all(1000 <= r.n <= 500_000 for r in a)
    assert all(r.steps % 100 == 0 and 500 <= r.steps <= 5000 for r in a)
    assert {r.kind for r in a} == {"drop", "hopper"}


def test_all_served_within_capacity():
    clock = serve.FakeClock()
    be = FakeBackend(clock)
    trace = serve.poisson_trace(seed=3, rate=20, duration=3, n_lo=1000, n_hi=50_000)
    s = _server(clock, be, max_particles=200_000)
    done = s.run(trace)
    assert len(done) == len(trace)
    assert be.max_particles <= 200_000
    assert all(r.finished >= r.req.arrival for r in done)
    assert all(r.admitted >= r.req.arrival for r in done)
    assert s.particle_steps == sum(r.req.steps * r.scene_np for r in done)
    rep = s.report(clock())
    assert rep["p99_latency_s"] >= rep["p50_latency_s"] > 0


def test_fcfs_and_overlap():
    """Requests that arrive while others run are admitted as soon as room frees
    up; several scenes run at once (continuous batching, not a queue)."""
    clock = serve.FakeClock()
    be = FakeBackend(clock)
    trace = serve.poisson_trace(seed=11, rate=50, duration=1, n_lo=1000, n_hi=20_000)
    s = _server(clock, be, max_particles=60_000)
    done = s.run(trace)
    assert be.max_live > 1
    adm = sorted(done, key=lambda r: r.req.arrival)
    t = [r.admitted for r in adm]
    assert t == sorted(t)  # first come, first served


def test_hbm_budget_defers_and_rejects():
    clock = serve.FakeClock()
    be = FakeBackend(clock)
    trace = [serve.Request(0, 0.0, "drop", 20_000, 200, 1), serve.Request(1, 0.0, "drop", 20_000, 200, 2),
             serve.Request(2, 0.0, "drop", 400_000, 200, 3)]
    sc0 = trace[0].scene()
    one = _server(clock, be).footprint(sc0)
    s = _server(clock, be, budget=int(one * 1.5))   # room for one small scene at a time
    done = s.run(trace)
    by = {r.req.rid: r for r in done}
    assert by[2].result == {"rejected": True}
    assert by[1].admitted >= by[0].finished   # deferred until the first one retired
    assert be.max_live == 1


def test_footprint_is_the_library_allocation_arithmetic():
    """Admission budgets with xdem_batch_device_bytes: linear in particles and
    cells, the history buffers added when a scene has history, and a budget of
    exactly two scenes' batch admits two of three equal scenes at a time."""
    b1 = serve.device_bytes(100_000, 1, 1 << 20, lists=False)
    b2 = serve.device_bytes(200_000, 1, 1 << 20, lists=False)
    b3 = serve.device_bytes(100_000, 1, 1 << 21, lists=False)
    per_p = (b2 - b1) / 100_000
    per_c = (b3 - b1) / (1 << 20)
    assert 200 < per_p < 400 and per_c == 8
    # candidate lists: 32 entries and a count per particle, a 16-bit tag per cell
    bl = serve.device_bytes(100_000, 1, 1 << 20)
    assert bl - b1 == 100_000 * 33 * 4 + 2 * (1 << 20)
    assert serve.device_bytes(100_000, 1, 1 << 20, history=True) - b1 >= 100_000 * 2 * 16 * 16
    clock = serve.FakeClock()
    be = FakeBackend(clock)
    trace = [serve.Request(k, 0.0, "drop", 20_000, 200, 1) for k in range(3)]
    sc = trace[0].scene()
    cells = int(np.prod(sc.grid_size))
    s = _server(clock, be, budget=serve.device_bytes(2 * sc.np, 2, 2 * cells))
    s.run(trace)
    assert be.max_live == 2
