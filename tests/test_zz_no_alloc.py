"""xdem_reorder_data_and_find_cell_start_noalloc (no allocation, fixed device
scratch) against xdem_reorder_data_and_find_cell_start (library workspace):
identical hash/index/sorted_index/cstart/cend, bit for bit, on two calls in a
row with the second smaller (the scratch keeps the first call's contents),
at the scratch's maximum size, and against the oracle's stable sort; a size
beyond the maximum is refused with nothing written."""
import numpy as np
import pytest

from metakonga_amd import abi, scenes
from xdem_testutil import DBuf

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20


def _keys(n, ncell, seed):
    u = scenes.splitmix64(seed, max(n, 1), 5)[:n]
    keys = np.sort((u % np.uint64(ncell)).astype(np.uint32))
    # nearly sorted, as between DEM steps, plus random displacements
    sw = (scenes.splitmix64(seed + 1, max(n // 50, 1), 6) % np.uint64(max(n, 1))).astype(np.int64)
    keys[sw] = (scenes.splitmix64(seed + 2, len(sw), 7) % np.uint64(ncell)).astype(np.uint32)
    return keys


def _run(lib, fn, keys, ncell):
    n = keys.size
    idx = np.arange(n, dtype=np.uint32)
    dh, di = DBuf(lib, keys), DBuf(lib, idx)
    dcs, dce, dso = DBuf(lib, nbytes=4 * ncell), DBuf(lib, nbytes=4 * ncell), DBuf(lib, nbytes=4 * n)
    dcs.set(np.full(ncell, 777, np.uint32))
    dce.set(np.full(ncell, 999, np.uint32))
    abi.check(lib, fn(dh.ptr, di.ptr, dcs.ptr, dce.ptr, dso.ptr, n, ncell))
    return (dh.get(np.uint32, n), di.get(np.uint32, n), dso.get(np.uint32, n),
            dcs.get(np.uint32, ncell), dce.get(np.uint32, ncell))


def test_noalloc_two_calls_match(lib, orc):
    cases = [(1_000_003, 1 << 21, 11), (70_001, 1 << 18, 12), (MAX_KEYS, 1 << 24, 13), (1, 8, 14)]
    for n, ncell, seed in cases:
        print(f"no-alloc reorder: n={n} ncell={ncell}", flush=True)
        keys = _keys(n, ncell, seed)
        a = _run(lib, lib.xdem_reorder_data_and_find_cell_start_noalloc, keys, ncell)
        b = _run(lib, lib.xdem_reorder_data_and_find_cell_start, keys, ncell)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (n, ncell)
        h, i2, cs, ce, srt = orc.reorder(keys, np.arange(n, dtype=np.uint32), ncell)
        for x, y in zip(a, (h, i2, srt, cs, ce)):
            assert np.array_equal(x, y), (n, ncell)


def test_noalloc_refuses_beyond_max(lib):
    n, ncell = MAX_KEYS + 1, 1 << 10
    keys = np.zeros(n, np.uint32)
    dh, di = DBuf(lib, keys), DBuf(lib, np.arange(n, dtype=np.uint32))
    dcs, dce, dso = DBuf(lib, nbytes=4 * ncell), DBuf(lib, nbytes=4 * ncell), DBuf(lib, nbytes=4 * n)
    dcs.set(np.full(ncell, 5, np.uint32))
    rc = lib.xdem_reorder_data_and_find_cell_start_noalloc(dh.ptr, di.ptr, dcs.ptr, dce.ptr,
                                                           dso.ptr, n, ncell)
    assert rc != 0
    assert b"exceeds" in lib.xdem_last_error()
    assert np.all(dcs.get(np.uint32, ncell) == 5)
