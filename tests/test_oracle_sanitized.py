"""The oracle's own pins rerun on its AddressSanitizer + UndefinedBehavior-
Sanitizer build (oracle/Makefile target `san`): out-of-bounds reads or
writes, use after free and undefined arithmetic in the restatement abort the
child process and fail this test.  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "oracle", "liboracle_san.so")
PINS = ["test_oracle_grid.py", "test_oracle_forces.py", "test_oracle_cohesion.py",
        "test_oracle_drum.py", "test_oracle_history.py", "test_oracle_release.py",
        "test_oracle_setb.py"]


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.slow
def test_oracle_pins_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("libasan runtime not found")
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, XDEM_ORACLE_LIB=SAN,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "-m", "not gpu"] + [os.path.join(ROOT, "tests", p) for p in PINS],
                       env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error" not in tail, tail
