"""The device's hash-table build (k_ht_prep / k_ht_insert / k_ht_finish in
piro_amd/csrc/assemble.hip) inserts the distinct keys concurrently by rank:
each key carries the CSR position of its first occurrence and probes with
atomicMin, taking over slots that hold larger ranks.  These CPU tests run that
procedure under random thread interleavings and compare the table with the
oracle's sequential methods::sethash (src/methods.cpp:74-118), which is itself
pinned to the reference's own tables in test_oracle.py.  High loads, long
clusters, wrap-around at the table's end and repeated keys are covered.
"""
import numpy as np
import pytest

import oracle as po

EMPTY = 0xFFFFFFFF


def parallel_build(A: po.CSR, ts: int, rng) -> tuple[np.ndarray, np.ndarray]:
    n = A.n_rows
    nnz = int(A.row_ptr[-1])
    kv = np.full(nnz, -1, np.int64)
    vv = np.zeros(nnz, np.float32)
    # k_ht_prep: first occurrences, each with its row's last value
    for r in range(n):
        b, e = int(A.row_ptr[r]), int(A.row_ptr[r + 1])
        for k in range(b, e):
            c = int(A.col[k])
            if any(int(A.col[j]) == c for j in range(b, k)):
                continue
            last = k
            for j in range(k + 1, e):
                if int(A.col[j]) == c:
                    last = j
            kv[k] = r * n + c
            vv[k] = A.val[last]
    distinct = int((kv >= 0).sum())
    if distinct > ts:
        raise RuntimeError("hash table full")
    # k_ht_insert: one carrier per distinct key, stepped in random order;
    # each step is one atomicMin on the carrier's current slot
    slots = np.full(ts, EMPTY, np.uint64)
    carriers = [[int(k), int(kv[k]) % ts] for k in np.nonzero(kv >= 0)[0]]
    occ = []
    while carriers:
        i = int(rng.integers(len(carriers)))
        t, p = carriers[i]
        old = int(slots[p])
        slots[p] = min(old, t)
        if old == EMPTY:
            occ.append(p)
            carriers.pop(i)
            continue
        if old > t:
            t = old
        carriers[i] = [t, (p + 1) % ts]
    # k_ht_finish
    keys = np.full(ts, -1, np.int32)
    vals = np.zeros(ts, np.float32)
    for p in occ:
        t = int(slots[p])
        keys[p] = kv[t]
        vals[p] = vv[t]
    return keys, vals


def random_csr(n, max_row, n_cols, rng):
    lens = rng.integers(0, max_row + 1, n)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, n_cols, int(rp[-1])).astype(np.int32)
    val = rng.standard_normal(int(rp[-1])).astype(np.float32)
    return po.CSR(rp, col, val)


@pytest.mark.parametrize("scheme", [po.UW, po.DW, po.CD])
@pytest.mark.parametrize("load", [0.3, 0.9, 1.0])
def test_rank_insert_equals_sequential_laplacian(scheme, load):
    rng = np.random.default_rng(int(load * 10) + scheme)
    A = po.lap_csr(5, 4, 3, scheme, np.float32(0.25))
    # distinct keys: the schemes with repeated self entries have fewer
    distinct = len({(r, int(c)) for r in range(A.n_rows)
                    for c in A.col[A.row_ptr[r]:A.row_ptr[r + 1]]})
    ts = max(1, int(distinct / load))
    rk, rv = po.ht_build(A, ts)
    for _ in range(3):
        keys, vals = parallel_build(A, ts, rng)
        np.testing.assert_array_equal(keys, rk)
        assert vals.tobytes() == rv.tobytes()


@pytest.mark.parametrize("seed", range(6))
def test_rank_insert_equals_sequential_random(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 40))
    # narrow column ranges make repeated keys inside rows
    A = random_csr(n, 9, int(rng.integers(1, n + 1)), rng)
    distinct = len({(r, int(c)) for r in range(n) for c in A.col[A.row_ptr[r]:A.row_ptr[r + 1]]})
    for ts in (distinct, distinct + 1, distinct + 7, 3 * distinct + 1):
        if ts < 1:
            continue
        rk, rv = po.ht_build(A, ts)
        keys, vals = parallel_build(A, ts, rng)
        np.testing.assert_array_equal(keys, rk)
        assert vals.tobytes() == rv.tobytes()


def test_rank_insert_wraps_around_the_table_end():
    # every key hashes to the last slots, so the clusters wrap to slot 0
    n = 6
    ts = 11
    rows, cols = [], []
    for r in range(n):
        for c in range(n):
            if (r * n + c) % ts >= ts - 3:
                rows.append(r)
                cols.append(c)
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    A = po.CSR(rp, np.array(cols, np.int32), np.arange(len(cols), dtype=np.float32))
    assert len(cols) <= ts
    rk, rv = po.ht_build(A, ts)
    assert rk[0] != -1  # a cluster crossed the end
    keys, vals = parallel_build(A, ts, np.random.default_rng(7))
    np.testing.assert_array_equal(keys, rk)
    assert vals.tobytes() == rv.tobytes()


def test_more_keys_than_slots_is_full():
    A = po.lap_csr(3, 3, 3, po.CD, np.float32(1.0))
    with pytest.raises(RuntimeError):
        po.ht_build(A, 20)
    with pytest.raises(RuntimeError):
        parallel_build(A, 20, np.random.default_rng(0))
