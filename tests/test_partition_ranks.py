"""The neighbour search's quicksort identity (CPU): the reference's
apartition / aquickSort (solver.cpp:600-632) restated with its two scans,
against the formulation the wave-per-node kernel uses — every partition from
ranks, disjoint ranges in any order, and the final sort stopped at the output
positions.

Partition of [lo, hi] with pivot p = a[lo]: let g_1 < g_2 < ... be the
positions in (lo, hi] holding a value > p and l_1 > l_2 > ... those holding a
value <= p, lo last.  The scans swap g_k and l_k for every k with g_k < l_k
(a prefix, K swaps) and the pivot lands at max(l_{K+1}, g_K), or l_1 if K = 0.
"""
import numpy as np
import pytest


def qpart_scans(wd, wm, l, r):
    pivot = wd[l]
    i, j = l, r + 1
    while True:
        i += 1
        while i <= r and wd[i] <= pivot:
            i += 1
        j -= 1
        while wd[j] > pivot:
            j -= 1
        if i >= j:
            break
        wd[i], wd[j] = wd[j], wd[i]
        wm[i], wm[j] = wm[j], wm[i]
    wd[l], wd[j] = wd[j], wd[l]
    wm[l], wm[j] = wm[j], wm[l]
    return j


def qsort_scans(wd, wm, l, r):
    """aquickSort with an explicit stack, the larger side pushed."""
    st = [(l, r)]
    while st:
        lo, hi = st.pop()
        while lo < hi:
            j = qpart_scans(wd, wm, lo, hi)
            if j - lo < hi - j:
                st.append((j + 1, hi))
                hi = j - 1
            else:
                st.append((lo, j - 1))
                lo = j + 1


def qpart_ranks(wd, wm, l, r):
    pivot = wd[l]
    pos = np.arange(l + 1, r + 1)
    gt = wd[l + 1:r + 1] > pivot
    g = pos[gt]
    lpos = np.concatenate([[l], pos[~gt]])[::-1]   # l_1, l_2, ..., lo last
    m = min(len(g), len(lpos))
    ok = g[:m] < lpos[:m]
    K = m if ok.all() else int(np.argmin(ok))
    jstar = int(max(lpos[K], g[K - 1])) if K > 0 else int(lpos[0])
    src = np.arange(len(wd))
    for k in range(K):
        src[g[k]], src[lpos[k]] = lpos[k], g[k]
    src[l], src[jstar] = src[jstar], src[l]
    wd[:] = wd[src]
    wm[:] = wm[src]
    return jstar


def qsort_ranks(wd, wm, l, r, need=None, small=0, rng=None):
    """Ranges taken in an arbitrary order; ranges of at most `small` elements
    sorted with the scans (as the kernel's single lanes do); ranges lying at
    or above `need` skipped."""
    need = r + 1 if need is None else need
    pending, shorts = [(l, r)], []
    while pending:
        k = int(rng.integers(len(pending))) if rng is not None else 0
        lo, hi = pending.pop(k)
        if lo >= hi or lo >= need:
            continue
        if hi - lo + 1 <= small:
            shorts.append((lo, hi))
            continue
        j = qpart_ranks(wd, wm, lo, hi)
        pending.append((lo, j - 1))
        if j + 1 < need:
            pending.append((j + 1, hi))
    for lo, hi in shorts[::-1]:
        qsort_scans(wd, wm, lo, hi)


def _arrays(seed, count):
    rng = np.random.default_rng(seed)
    for t in range(count):
        n = int(rng.integers(1, 90))
        kind = t % 4
        if kind == 0:
            wd = rng.integers(0, 3, n).astype(float)
        elif kind == 1:
            wd = rng.integers(0, 10, n).astype(float)
        elif kind == 2:
            wd = rng.random(n)
        else:  # penalised keys: few distances, 1e4^rank factors
            wd = np.round(rng.random(n) * 4) * 1e4 ** rng.integers(0, 3, n)
        yield rng, wd


def test_partition_by_ranks_equals_the_scans():
    for rng, wd in _arrays(1, 3000):
        n = len(wd)
        if n < 2:
            continue
        a, b = wd.copy(), np.arange(n)
        ja = qpart_scans(a, b, 0, n - 1)
        c, d = wd.copy(), np.arange(n)
        jc = qpart_ranks(c, d, 0, n - 1)
        assert ja == jc and np.array_equal(b, d)


@pytest.mark.parametrize("small", [0, 8])
def test_quicksort_by_ranks_equals_the_reference(small):
    for rng, wd in _arrays(2, 3000):
        n = len(wd)
        a, b = wd.copy(), np.arange(n)
        qsort_scans(a, b, 0, n - 1)
        c, d = wd.copy(), np.arange(n)
        qsort_ranks(c, d, 0, n - 1, small=small, rng=rng)
        assert np.array_equal(b, d)


def test_sort_stopped_at_the_output_positions():
    for rng, wd in _arrays(3, 3000):
        n = len(wd)
        need = int(rng.integers(1, n + 1))
        a, b = wd.copy(), np.arange(n)
        qsort_scans(a, b, 0, n - 1)
        c, d = wd.copy(), np.arange(n)
        qsort_ranks(c, d, 0, n - 1, need=need, small=8, rng=rng)
        assert np.array_equal(b[:need], d[:need])


def test_a_plausible_slip_is_caught():
    """Landing the pivot at l_{K+1} alone (forgetting that the j-scan can stop
    on a position the swaps filled) breaks the identity on these inputs."""
    bad = 0
    for rng, wd in _arrays(4, 500):
        n = len(wd)
        if n < 3:
            continue
        pivot = wd[0]
        pos = np.arange(1, n)
        gt = wd[1:] > pivot
        g = pos[gt]
        lpos = np.concatenate([[0], pos[~gt]])[::-1]
        m = min(len(g), len(lpos))
        ok = g[:m] < lpos[:m]
        K = m if ok.all() else int(np.argmin(ok))
        a, b = wd.copy(), np.arange(n)
        j = qpart_scans(a, b, 0, n - 1)
        if K > 0 and j != int(lpos[K]):
            bad += 1
    assert bad > 0


def codes_of(wd):
    """The sort kernel's keys: for each entry, the number of entries strictly
    below it."""
    return np.array([(wd < v).sum() for v in wd], dtype=np.int64)


def test_quicksort_on_codes_equals_quicksort_on_keys():
    """The final quicksort runs on the codes instead of the penalised keys
    (one lane per node, one byte per key): the permutation of the payload must
    be the one the reference's quicksort gives on the keys themselves, ties
    and the stop at the output positions included."""
    for rng, wd in _arrays(5, 3000):
        n = len(wd)
        need = int(rng.integers(1, n + 1))
        a, b = wd.copy(), np.arange(n)
        qsort_scans(a, b, 0, n - 1)
        cd = codes_of(wd)
        assert cd.max() < 256
        c, d = cd.astype(float), np.arange(n)
        qsort_ranks(c, d, 0, n - 1, need=need, small=n, rng=rng)  # all by the scans, stopped
        assert np.array_equal(b[:need], d[:need])


def test_codes_that_merge_distinct_keys_are_caught():
    """A lossy code (here: keys rounded before they are ranked, which merges
    distinct keys) changes the permutation on these inputs, so the identity
    above is not vacuous."""
    bad = 0
    for rng, wd in _arrays(6, 2000):
        n = len(wd)
        a, b = wd.copy(), np.arange(n)
        qsort_scans(a, b, 0, n - 1)
        merged = np.round(wd / 4.0) * 4.0  # distinct keys merged
        c, d = merged.copy(), np.arange(n)
        qsort_scans(c, d, 0, n - 1)
        bad += int(not np.array_equal(b, d))
    assert bad > 0
