"""Pins of the oracle's grid hash, sort and cell ranges (CPU only).

References: calcGridPos/calcGridHash include/mphysics_cuda_impl.cuh:65-83,
grid_base::getCellNumber/getHash source/grid_base.cpp:139-155,
thrust::sort_by_key source/mphysics_cuda_def.cu:92-94 (stable radix sort),
reorderDataAndFindCellStart_kernel include/mphysics_cuda_impl.cuh:166-212.
"""
import numpy as np
import pytest

from metakonga_amd import abi, scenes


def params(wo=(0.0, 0.0, 0.0), cs=1.0, gs=(4, 4, 4)):
    d = abi.DeviceParameters()
    d.grid_size = abi.uint3(*gs)
    d.ncell = gs[0] * gs[1] * gs[2]
    d.cell_size = cs
    d.world_origin = abi.double3(*wo)
    d.dt = 1e-5
    d.half2dt = 0.5e-10
    return d


def test_hash_exhaustive_tiny_grid(orc):
    """Cell centres of every cell of a 4x8x2 grid, and copies shifted by whole
    grid periods (wrap by mask), map to z*gy*gx + y*gx + x."""
    gs = (4, 8, 2)
    p = params(wo=(-0.5, 1.25, 3.0), cs=0.25, gs=gs)
    for cz in range(-gs[2], 2 * gs[2]):
        for cy in range(-gs[1], 2 * gs[1]):
            for cx in range(-gs[0], 2 * gs[0]):
                x = -0.5 + (cx + 0.5) * 0.25
                y = 1.25 + (cy + 0.5) * 0.25
                z = 3.0 + (cz + 0.5) * 0.25
                c = orc.cell_of(p, x, y, z)
                assert tuple(c) == (cx, cy, cz)
                want = (cz % gs[2]) * gs[1] * gs[0] + (cy % gs[1]) * gs[0] + (cx % gs[0])
                assert orc.hash_of(p, c) == want


def test_hash_on_cell_faces(orc):
    """A point exactly on a face belongs to the upper cell (floor), both in
    positive and negative coordinates."""
    p = params(wo=(0.0, 0.0, 0.0), cs=0.5, gs=(8, 8, 8))
    for k in range(-6, 7):
        c = orc.cell_of(p, k * 0.5, 0.0, -k * 0.5)
        assert c[0] == k and c[2] == -k and c[1] == 0


def test_host_mode_abs_below_origin(orc):
    """The CPU path takes abs() of the floored cell (A6): differs only below wo."""
    p = params(wo=(0.0, 0.0, 0.0), cs=1.0, gs=(8, 8, 8))
    assert tuple(orc.cell_of(p, -2.5, 1.5, 3.5, orc.ORC_GPU)) == (-3, 1, 3)
    assert tuple(orc.cell_of(p, -2.5, 1.5, 3.5, orc.ORC_HOST)) == (3, 1, 3)
    assert tuple(orc.cell_of(p, 2.5, 1.5, 3.5, orc.ORC_HOST)) == (2, 1, 3)


def test_hash_and_index_kernel_layout(orc):
    p = params(wo=(-1, -1, -1), cs=0.01, gs=(128, 128, 128))
    pos = np.zeros((5, 4))
    pos[:, 0] = [-1.0, -0.995, 0.275, 0.28, 0.3]
    h, idx = orc.hash_and_index(p, pos)
    assert list(idx) == [0, 1, 2, 3, 4]
    # x cells 0, 0, 127, 128->0 (wrap), 130->2
    assert list(h % 128) == [0, 0, 127, 0, 2]
    h2, idx2 = orc.hash_and_index(p, pos[:2], off=5, hash_=np.zeros(7, np.uint32),
                                  index=np.zeros(7, np.uint32))
    assert list(idx2[5:]) == [5, 6]


@pytest.mark.parametrize("n,ncell,dup", [(1, 8, 1), (7, 8, 1), (1000, 64, 1), (5000, 1 << 21, 1),
                                         (3000, 4096, 50)])
def test_sort_and_cell_ranges(orc, n, ncell, dup):
    u = scenes.splitmix64(n + ncell, n, 5)
    keys = (u % np.uint64(max(ncell // dup, 1))).astype(np.uint32) * np.uint32(dup)
    keys = np.minimum(keys, ncell - 1).astype(np.uint32)
    idx = np.arange(n, dtype=np.uint32)
    h, i2, cs, ce, srt = orc.reorder(keys, idx, ncell)
    # stable sort by definition
    order = sorted(range(n), key=lambda j: int(keys[j]))
    assert list(i2) == order
    assert np.array_equal(h, keys[order])
    assert np.array_equal(srt, i2)
    # cell range invariants
    occupied = np.unique(keys)
    empty = np.setdiff1d(np.arange(ncell), occupied)
    assert np.all(cs[empty] == 0xFFFFFFFF) and np.all(ce[empty] == 0)
    assert int((ce[occupied].astype(np.int64) - cs[occupied]).sum()) == n
    for c in occupied[:200]:
        assert np.all(h[cs[c]:ce[c]] == c)
        assert cs[c] == 0 or h[cs[c] - 1] != c
        assert ce[c] == n or h[ce[c]] != c


def test_pair_list_equals_bruteforce(orc):
    """With cs >= 2 r_max and a non-aliasing grid, the 27-cell pair list is the
    all-pairs predicate set."""
    from xdem_testutil import dense_state
    pos, vel, om = dense_state(n_side=8, pitch_factor=1.97, jitter=0.05)
    r = pos[0, 3]
    p = params(wo=(0.0, -0.6, 0.0), cs=2 * r, gs=(32, 32, 32))
    p.np = pos.shape[0]
    h, idx = orc.hash_and_index(p, pos)
    h, idx, cs, ce, srt = orc.reorder(h, idx, p.ncell)
    n = pos.shape[0]
    m, i = scenes.mass_inertia(pos[:, 3], 2500.0)
    f = np.zeros((n, 3))
    t = np.zeros((n, 3))
    cp = scenes.contact_property(scenes.MaterialSpec(), scenes.MaterialSpec(), scenes.ContactSpec())
    grid_pairs = orc.p2p(p, 1, pos, vel, om, f, t, m, srt, cs, ce, cp, want_pairs=True)
    bf = orc.pairs_bruteforce(pos)
    assert len(grid_pairs) > n  # contact-rich
    assert set(map(tuple, grid_pairs.tolist())) == set(map(tuple, bf.tolist()))
    # the list is symmetric: (i,k) present iff (k,i) present
    s = set(map(tuple, bf.tolist()))
    assert all((k, i) in s for i, k in s)
