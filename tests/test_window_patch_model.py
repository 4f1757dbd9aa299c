"""The window-patch re-sort (k_patch) restated on the CPU and checked
against a full stable re-sort: movers ranked by (key, id), their storage
indices ascending, windows [src, lo - 1] / [lo, src] merged where they overlap
or touch, every window position's element found by binary search (the mover
placed there, or the stayer of rank q at q + #{i : sAsc[i] - i <= q}), cell
boundaries rewritten at window positions, emptied cells cleared and the
previous windows reset to the identity.  Ranks are worked out as the kernel's
block_rank does (per chunk by counting, across chunks by binary search), a
position's cell boundary takes the key before it from the slot before, and a
cell is emptied when every position of it holds a mover and none enters.
Random cases: the permutation, the keys at every position and the cell ranges
must equal the full re-sort's.
This pins the algorithm the GPU kernel implements; the GPU tests compare the
kernel with the full merge and with the oracle's sort.
"""
import bisect
import numpy as np

EMPTY = 0xFFFFFFFF


def ranges(skeys, ncell):
    cr = np.full((ncell, 2), [EMPTY, 0], dtype=np.int64)
    n = len(skeys)
    for j in range(n):
        if j == 0 or skeys[j - 1] != skeys[j]:
            cr[skeys[j], 0] = j
            if j > 0:
                cr[skeys[j - 1], 1] = j
    if n:
        cr[skeys[n - 1], 1] = n
    return cr


def count_lt(arr, x):  # sorted arr
    return bisect.bisect_left(arr, x)


def count_le(arr, x):
    return bisect.bisect_right(arr, x)


def block_rank(keys, ch=64):
    """Rank of each of distinct keys among all, as the kernel's block_rank:
    counting inside the key's chunk of ch, binary search over the other
    chunks in key order (no tie is tested: the keys must be distinct)."""
    m = len(keys)
    assert len(set(keys)) == m
    chunks = [sorted(keys[c:c + ch]) for c in range(0, m, ch)]
    out = []
    for t in range(m):
        w, k = t // ch, keys[t]
        r = sum(1 for i in range(w * ch, min(m, w * ch + ch)) if keys[i] < k)
        for c, srt in enumerate(chunks):
            if c != w:
                r += bisect.bisect_left(srt, k)
        out.append(r)
    return out


def patch(keys, ids, newkeys, movers, cr_old, prev_windows, perm_in, ch=64):
    """keys/ids: old sorted order (storage); movers: storage indices (append
    order); returns perm, skey, crange, windows — mirroring the kernel."""
    n = len(keys)
    M = len(movers)
    rX = [(int(newkeys[s]) << 32) | int(ids[s]) for s in movers]
    rS = [int(s) for s in movers]
    oldX = [(int(keys[q]) << 32) | int(ids[q]) for q in range(n)]
    # insertion point recorded with the mover: previous entries below X
    rL = [bisect.bisect_left(oldX, x) for x in rX]
    # P1: ranks by (key, id) and by storage index (block_rank)
    sX = [0] * M; sSrc = [0] * M; sOut = [0] * M; sAsc = [0] * M
    wA = [0] * M; wB = [0] * M
    rks, rss = block_rank(rX, ch), block_rank(rS, ch)
    for t in range(M):
        sX[rks[t]] = rX[t]; sSrc[rks[t]] = rS[t]
        sAsc[rss[t]] = rS[t]
    for t in range(M):
        rk = rks[t]
        below = count_lt(sAsc, rL[t])
        sOut[rk] = rk + rL[t] - below
        wA[t] = rS[t] if rS[t] < rL[t] else rL[t]
        wB[t] = rL[t] - 1 if rS[t] < rL[t] else rS[t]
    # P2: windows ranked by (a, append index) (block_rank; distinct keys)
    sWa = [0] * M; sWb = [0] * M
    rws = block_rank([(wA[t] << 32) | t for t in range(M)], ch)
    for t in range(M):
        sWa[rws[t]] = wA[t]; sWb[rws[t]] = wB[t]
    # P3: merge by running max; compaction
    run = []
    m = -1
    for t in range(M):
        m = max(m, sWb[t])
        run.append(m)
    starts = [t == 0 or sWa[t] > run[t - 1] + 1 for t in range(M)]
    win = []
    for t in range(M):
        if starts[t]:
            win.append([sWa[t], None])
        if t + 1 == M or starts[t + 1]:
            win[-1][1] = run[t]
    nw = len(win)
    wl = [w[1] - w[0] + 2 for w in win]
    wpre = [0]
    for x in wl:
        wpre.append(wpre[-1] + x)
    total_new = wpre[-1]
    pl = [b - a + 1 for a, b in prev_windows]
    ppre = [0]
    for x in pl:
        ppre.append(ppre[-1] + x)
    total = total_new + ppre[-1]
    starts_sorted = [w[0] for w in win]

    def elem(j):
        c1 = count_lt(sOut, j)
        if c1 < M and sOut[c1] == j:
            return sSrc[c1], int(sX[c1] >> 32)
        q0 = j - c1
        lo, hi = 0, M
        while lo < hi:
            mid = (lo + hi) // 2
            if sAsc[mid] - mid <= q0:
                lo = mid + 1
            else:
                hi = mid
        s = min(q0 + lo, n - 1)
        return s, int(keys[s])

    def in_new(j):
        k = bisect.bisect_right(starts_sorted, j) - 1
        return k >= 0 and j <= win[k][1]

    perm = perm_in.copy()
    skey = newkeys.copy()
    cr = cr_old.copy()
    # the key each new-window slot works out (the kernel's lanes pass it on
    # to the lane of the next slot)
    kcs = [None] * total_new
    for g in range(total_new):
        k = bisect.bisect_right(wpre, g) - 1
        a, b = win[k]
        j = a + (g - wpre[k])
        if j <= b:
            kcs[g] = elem(j)[1]
        elif j < n:
            kcs[g] = int(keys[j])
    # writes collected per position, then applied (the kernel's threads write
    # disjoint positions; boundaries are written once each)
    for g in range(total):
        if g < total_new:
            k = bisect.bisect_right(wpre, g) - 1
            a, b = win[k]
            j = a + (g - wpre[k])
            if j > n:
                continue
            kp = None
            kc = kcs[g]
            if j <= b:
                perm[j] = elem(j)[0]
                skey[j] = kc
            if j > 0:
                kp = kcs[g - 1] if j - 1 >= a else int(keys[j - 1])
            if j == 0 or j == n or kp != kc:
                if j < n:
                    cr[kc, 0] = j
                if j > 0:
                    cr[kp, 1] = j
        else:
            gg = g - total_new
            k = bisect.bisect_right(ppre, gg) - 1
            j = prev_windows[k][0] + (gg - ppre[k])
            if j < n and not in_new(j):
                perm[j] = j
    # emptied cells: every position of the old cell holds a mover, none enters
    mark = set(rS)
    klist = [int(x >> 32) for x in sX]
    for t in range(M):
        ko = int(keys[rS[t]])
        first = int(cr_old[ko, 0])
        size = int(cr_old[ko, 1]) - first
        if all(q in mark for q in range(first, first + size)) and \
                count_le(klist, ko) == count_lt(klist, ko):
            cr[ko] = [EMPTY, 0]
    return perm, skey, cr, [tuple(w) for w in win]


def trial(rng, n, ncell, M, prev, ch=64):
    ids = rng.permutation(n)
    keys = np.sort(rng.integers(0, ncell, n))
    o = np.lexsort((ids, keys))
    keys = keys[o]
    ids = ids[o]
    cr_old = ranges(keys, ncell)
    movers = rng.choice(n, size=M, replace=False)
    newkeys = keys.copy()
    for s in movers:
        k = rng.integers(0, ncell)
        while k == keys[s]:
            k = rng.integers(0, ncell)
        newkeys[s] = k
    perm_ref = np.lexsort((ids, newkeys))
    # the previous sort's windows: perm holds garbage there (a permutation of
    # their own positions), the identity elsewhere
    perm_in = np.arange(n)
    for a, b in prev:
        perm_in[a:b + 1] = rng.permutation(np.arange(a, b + 1))
    perm, skey, cr, win = patch(keys, ids, newkeys, movers, cr_old, prev, perm_in, ch)
    assert np.array_equal(perm, perm_ref), (n, M)
    # keys at every position: inside windows as written, outside unchanged
    want_keys = newkeys[perm_ref]
    inU = np.zeros(n, bool)
    for a, b in win:
        inU[a:b + 1] = True
    assert np.array_equal(np.where(inU, skey, keys), want_keys), (n, M)
    ref = ranges(want_keys, ncell)
    assert np.array_equal(cr, ref), (n, M)
    return win




def test_block_rank_is_the_rank_of_distinct_keys():
    rng = np.random.default_rng(3)
    for it in range(300):
        m = int(rng.integers(0, 200))
        # distinct keys, clustered (a small first part and the index below it)
        keys = [(int(x) << 32) | t for t, x in enumerate(rng.integers(0, int(rng.integers(1, 40)), m))]
        want = sorted(range(m), key=lambda t: keys[t])
        for ch in (1, 5, 64):
            got = block_rank(keys, ch)
            assert sorted(got) == list(range(m))
            assert [t for _, t in sorted(zip(got, range(m)))] == want


def test_window_patch_matches_full_resort():
    rng = np.random.default_rng(5)
    for it in range(1500):
        n = int(rng.integers(1, 120))
        ncell = int(rng.integers(2, 50))
        M = int(rng.integers(0, min(n, 100) + 1))
        prev = []
        p = 0
        while p < n and len(prev) < 5:
            a = int(rng.integers(p, n))
            b = int(rng.integers(a, min(n, a + 10)))
            prev.append((a, b))
            p = b + 2
        # chunks of 64 as on the GPU, and of 7 (many chunks)
        trial(rng, n, ncell, M, prev, 64 if it % 2 else 7)
