"""State files (CPU): the reference's part%04d.bin and .bfr byte layouts
(dem_simulation::saveResult / saveFinalResult, dem_simulation.cpp:491-583;
xDynamicsSolver::saveFinalResult, xDynamicsSolver.cpp:111-130) written by the
library against the same records packed here with struct, read back in both
precisions, and the checkpoint container's rejection of damaged files.  The
container hash is FNV-1a 64 (offset basis 14695981039346656037, prime
1099511628211), restated here independently."""
import struct

import numpy as np
import pytest

from metakonga_amd import abi, state_io


def _fnv1a(b: bytes) -> int:
    h = 14695981039346656037
    for c in b:
        h ^= c
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.fixture()
def rng():
    return np.random.default_rng(4)


def test_part_f64_layout_and_read(tmp_path, rng):
    pos = rng.normal(size=(37, 4))
    p = tmp_path / "part0003.bin"
    state_io.write_part(p, 0.125, pos)
    assert p.read_bytes() == struct.pack("<dI", 0.125, 37) + pos.astype("<f8").tobytes()
    ct, back, prec = state_io.read_part(p)
    assert ct == 0.125 and prec == 2 and np.array_equal(back, pos)


def test_part_f32_layout_and_read(tmp_path, rng):
    pos = rng.normal(size=(11, 4)).astype(np.float32)
    p = tmp_path / "part0000.bin"
    state_io.write_part(p, 2.5, pos, precision=1)
    assert p.read_bytes() == struct.pack("<dI", 2.5, 11) + pos.astype("<f4").tobytes()
    ct, back, prec = state_io.read_part(p)
    assert prec == 1 and np.array_equal(back, pos.astype(np.float64))


def test_part_rejects_bad_sizes(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(struct.pack("<dI", 0.0, 5) + b"\0" * (32 * 5 - 1))
    with pytest.raises(RuntimeError):
        state_io.read_part(p)
    p.write_bytes(b"\0" * 7)
    with pytest.raises(RuntimeError):
        state_io.read_part(p)


@pytest.mark.parametrize("precision", [2, 1])
def test_bfr_layout_and_read(tmp_path, rng, precision):
    n = 23
    dt = np.float64 if precision == 2 else np.float32
    pos, vel, av = (rng.normal(size=(n, k)).astype(dt) for k in (4, 3, 3))
    p = tmp_path / "m_final.bfr"
    state_io.write_bfr(p, 0.75, pos, vel, av, precision=precision)
    tag = "<f8" if precision == 2 else "<f4"
    want = struct.pack("<dii", 0.75, precision, 1) + b"".join(
        a.astype(tag).tobytes() for a in (pos, vel, av))
    assert p.read_bytes() == want
    ct, P, V, A, prec = state_io.read_bfr(p, n)
    assert ct == 0.75 and prec == precision
    for got, ref in ((P, pos), (V, vel), (A, av)):
        assert np.array_equal(got, ref.astype(np.float64))
    # the reader takes np from the model: fewer is fine, more is refused
    state_io.read_bfr(p, n - 3)
    with pytest.raises(RuntimeError):
        state_io.read_bfr(p, n + 1)


def test_bfr_rejects_flag_and_precision(tmp_path):
    p = tmp_path / "x.bfr"
    p.write_bytes(struct.pack("<dii", 0.0, 2, 0) + b"\0" * 80)
    with pytest.raises(RuntimeError):
        state_io.read_bfr(p, 1)
    p.write_bytes(struct.pack("<dii", 0.0, 3, 1) + b"\0" * 80)
    with pytest.raises(RuntimeError):
        state_io.read_bfr(p, 1)


def _container(kind, payload, version=1, magic=b"XDEMCKP1"):
    return (magic + struct.pack("<IIQ", kind, version, len(payload)) + payload
            + struct.pack("<Q", _fnv1a(payload)))


@pytest.mark.parametrize("kind", [1, 2])
def test_checkpoint_container(tmp_path, kind):
    payload = bytes(range(256)) * 3
    p = tmp_path / "c.ckpt"
    p.write_bytes(_container(kind, payload))
    assert state_io.verify_checkpoint(p) == kind
    good = p.read_bytes()
    damaged = {
        "truncated": good[:-9],
        "flipped": good[:40] + bytes([good[40] ^ 1]) + good[41:],
        "magic": _container(kind, payload, magic=b"XDEMCKP0"),
        "version": _container(kind, payload, version=2),
        "kind": _container(3, payload),
        "empty": b"",
    }
    for name, data in damaged.items():
        p.write_bytes(data)
        with pytest.raises(RuntimeError):
            state_io.verify_checkpoint(p)
        assert name


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(RuntimeError):
        state_io.read_part(tmp_path / "nope.bin")
    lib = abi.load()
    assert "nope" in lib.xdem_last_error().decode()
