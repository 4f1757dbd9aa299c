"""Build recipe of libxdem.so (hipcc, gfx950) — invoked by __graft_entry__.build().

Flags:
  --offload-arch=gfx950          MI355X (CDNA4)
  -ffp-contract=off              no FMA contraction: expressions round as written,
                                 as in the CPU oracle (bit-level parity of FP64)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                 no packed FP32 instructions (v_pk_*_f32)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libxdem.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["xdem_runtime.hip", "xdem_sort.hip", "xdem_f64.hip", "xdem_host.hip",
           "xdem_fused.hip", "xdem_batch.hip", "xdem_setb.hip", "xdem_io.hip"]

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
         "-I" + os.path.join(ROOT, "include")]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = False, jobs: int = 6) -> str:
    """Compile every source to an object (in parallel) and link libxdem.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            procs[[p.poll() is None for _, p in procs].index(True)][1].wait()
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for src, out in failed:
            sys.stderr.write(f"=== {src}\n{out}\n")
        raise RuntimeError("hipcc failed: " + ", ".join(os.path.basename(s) for s, _ in failed))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
