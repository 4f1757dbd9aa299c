"""The block log's records and JSON lines, on a stand-in batch (no GPU)."""
import io
import json

from metakonga_amd import blocklog


class _FakeBatch:
    def __init__(self):
        self.sizes = {0: 100, 1: 50}
        self.steps = 0
        self.prepared = 0

    def sync(self):
        pass

    def step(self, k):
        self.steps += k

    def prepare(self):
        self.prepared += 1

    def contact_metrics(self, sid):
        return (3 * self.sizes[sid], 1e-5 * (sid + 1))

    def energy(self, sid):
        return (0.5 * self.steps, (1.0, 2.0, float(sid)))


def test_block_lines_and_log():
    b = _FakeBatch()
    buf = io.StringIO()
    recs = blocklog.run_logged(b, [0, 1], 25, 10, buf)
    lines = [json.loads(x) for x in buf.getvalue().splitlines()]
    assert lines == recs
    assert len(lines) == 6 and b.prepared == 3 and b.steps == 25
    assert [r["steps"] for r in lines] == [10, 10, 10, 10, 5, 5]
    assert [r["block_end_step"] for r in lines] == [10, 10, 20, 20, 25, 25]
    r = lines[1]
    assert r["scene"] == 1 and r["particles"] == 50
    assert r["contacts_per_particle"] == 6.0 and r["contact_pairs"] == 150
    assert r["max_overlap"] == 2e-5 and r["momentum"] == [1.0, 2.0, 1.0]
    assert r["hbm_bytes_algorithmic"] == 192 * 150 * 10
    assert r["particle_steps_per_s"] > 0
