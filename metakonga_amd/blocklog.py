"""Per-block JSON metrics of a batch run.

After each step block one JSON line per scene: the block's whole-batch
particle-steps/s, and for the scene its contacts per particle, largest
overlap, kinetic energy and momentum at the block's end, plus the block's
algorithmic HBM bytes (192 B per particle-step on the fused path).  The
metrics are taken outside the timed block (a prepare() re-sorts the storage
and restarts the candidate lists, which changes no result).
"""
import json
import time

FP32C_BYTES_PER_PARTICLE_STEP = 192


def block_lines(b, sids, end_step, steps, seconds, n_total):
    """The JSON-ready records of one finished block (after b.prepare())."""
    out = []
    for sid in sids:
        n = b.sizes[sid]
        pairs, ov = b.contact_metrics(sid)
        ke, mom = b.energy(sid)
        out.append({
            "block_end_step": int(end_step),
            "steps": int(steps),
            "scene": int(sid),
            "particles": int(n),
            "particle_steps_per_s": n_total * steps / seconds if seconds > 0 else None,
            "contacts_per_particle": 2.0 * pairs / n if n else 0.0,
            "contact_pairs": int(pairs),
            "max_overlap": float(ov),
            "kinetic_energy": float(ke),
            "momentum": [float(x) for x in mom],
            "hbm_bytes_algorithmic": FP32C_BYTES_PER_PARTICLE_STEP * n_total * steps,
        })
    return out


def run_logged(b, sids, steps, block, out):
    """Advance the batch `steps` steps in blocks of `block`, writing one JSON line
    per block and scene to the text stream `out`.  Returns the records."""
    n_total = sum(b.sizes[s] for s in sids)
    done = 0
    recs = []
    while done < steps:
        k = min(block, steps - done)
        b.sync()
        t0 = time.perf_counter()
        b.step(k)
        b.sync()
        dt = time.perf_counter() - t0
        done += k
        b.prepare()
        for r in block_lines(b, sids, done, k, dt, n_total):
            out.write(json.dumps(r) + "\n")
            recs.append(r)
        out.flush()
    return recs
