"""CPU model of the kd builder's partition schedule (csrc/knn_build.hip:
kd_flag_kernel, kd_count_kernel, kd_before_kernel, kd_scatter_kernel and the
level loop of build_kd_ws).

A global kd level only reads its split list, so a list waits until its own
dimension splits (or until it has waited two levels) and is then moved down
the k <= 2 levels it waited by one stable partition on a 1- or 2-bit code.
This model restates that index arithmetic in numpy -- run codes from the
pair map, bucket-major offsets of the scanned tile counts, run sizes from
the deterministic split sizes, a position's segment found by descending
from the root -- and checks, for several split-list counts, that every
partition is a permutation inside the list and that the lists equal the
ones a partition of every list at every level gives (the rule of
src/KdTreeGpu.cuh:231-238: the left child is the first ceil(n/2) entries of
its segment in the split list).  The GPU build itself is checked against
the oracle by tests/test_gpu_parity.py."""
import numpy as np
import pytest


def _ordered(x):
    b = x.view(np.uint32).astype(np.uint64)
    return np.where(b >> 31, (~b) & 0xFFFFFFFF, b | 0x80000000)


def _tables(n, levels):
    st, sz = np.array([0]), np.array([n])
    for _ in range(levels):
        left = (sz + 1) >> 1
        st = np.stack([st, st + left], 1).ravel()
        sz = np.stack([left, sz - left], 1).ravel()
    return st, sz


def _descend(j, n, levels):
    st = np.zeros_like(j)
    sz = np.full_like(j, n)
    seg = np.zeros_like(j)
    for _ in range(levels):
        left = (sz + 1) >> 1
        r = (j - st) >= left
        seg = 2 * seg + r
        st = np.where(r, st + left, st)
        sz = np.where(r, sz - left, left)
    return st, sz, seg


def _flush(lst, code_map, n, level, k):
    """One partition of `lst` (segmented at level - k) by the pair map of
    level - 1, as kd_count_kernel / kd_before_kernel / kd_scatter_kernel."""
    code = code_map[lst]
    c = 3 - code if k == 2 else 1 - (code & 1)
    totals = np.array([(c == b).sum() for b in range(4)])
    offsets = np.concatenate([[0], np.cumsum(totals)[:-1]])
    g = np.zeros(n, np.int64)
    for b in range(4):
        m = c == b
        g[m] = offsets[b] + np.arange(m.sum())
    _, sz_l = _tables(n, level)
    nseg = (1 << level) >> k
    runs = np.array([sz_l[(s << k) + cc] for cc in range(1 << k) for s in range(nseg)])
    before = np.concatenate([[0], np.cumsum(runs)[:-1]])
    st, sz, seg = _descend(np.arange(n), n, level - k)
    for b in range(k - 1, -1, -1):
        left = (sz + 1) >> 1
        r = ((c >> b) & 1) == 1
        st = np.where(r, st + left, st)
        sz = np.where(r, sz - left, left)
    dst = st + (g - before[c * nseg + seg])
    assert dst.min() >= 0 and dst.max() < n
    assert len(np.unique(dst)) == n
    out = np.empty(n, np.int64)
    out[dst] = lst
    return out


@pytest.mark.parametrize("n,sd,levels,seed", [(100003, 3, 8, 0), (4097, 3, 9, 1), (77777, 2, 7, 2),
                                              (50000, 5, 9, 3), (65536, 1, 6, 4), (30011, 4, 10, 5)])
def test_lazy_partitions_equal_eager_partitions(n, sd, levels, seed):
    rng = np.random.default_rng(seed)
    pts = rng.standard_normal((n, sd)).astype(np.float32)
    pts[rng.integers(0, n, n // 10)] = 0.5  # ties
    keys = [_ordered(pts[:, d].copy()) for d in range(sd)]
    ids = np.arange(n, dtype=np.uint64)
    lazy = [np.lexsort((ids, keys[d])).astype(np.int64) for d in range(sd)]
    eager = [p.copy() for p in lazy]
    seg_by_id = np.zeros(n, np.int64)
    pend = [0] * sd
    maps = [None, None]
    for depth in range(levels):
        dim = depth % sd
        for j in range(sd):
            if pend[j] == 2 or (j == dim and pend[j] > 0):
                lazy[j] = _flush(lazy[j], maps[(depth - 1) & 1], n, depth, pend[j])
                pend[j] = 0
        assert np.array_equal(lazy[dim], eager[dim])
        st_l, sz_l = _tables(n, depth)
        piv = lazy[dim][st_l + ((sz_l + 1) >> 1) - 1]
        key = keys[dim]
        s = seg_by_id
        left = (key < key[piv][s]) | ((key == key[piv][s]) & (np.arange(n) <= piv[s]))
        # The eager schedule: every other list partitioned at every level.
        stp, _, segp = _descend(np.arange(n), n, depth)
        for j in range(sd):
            if j == dim:
                continue
            f = left[eager[j]]
            out = np.empty(n, np.int64)
            for q in range(1 << depth):
                m = segp == q
                seg = eager[j][m]
                fl = f[m]
                out[st_l[q]:st_l[q] + sz_l[q]] = np.concatenate([seg[fl], seg[~fl]])
            eager[j] = out
        prev = (maps[(depth - 1) & 1] & 1) if depth else np.zeros(n, np.int64)
        maps[depth & 1] = (prev << 1) | left.astype(np.int64)
        seg_by_id = 2 * s + (~left)
        for j in range(sd):
            if j != dim:
                pend[j] += 1
    for j in range(sd):
        if pend[j] > 0:
            lazy[j] = _flush(lazy[j], maps[(levels - 1) & 1], n, levels, pend[j])
        assert np.array_equal(lazy[j], eager[j])
