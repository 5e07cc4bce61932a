"""Host-side instance set-up follows the reference's semantics (CSR, cluster cap, canonical triangles, levels)."""
import numpy as np

from lagbound_amd import instances as I


def test_csr_symmetric_sorted():
    n, keys = I.chung_lu(3000, 2.1, 10.0, 1)
    rowptr, col, w = I.csr_from_keys(n, keys)
    assert w is None and rowptr[0] == 0 and rowptr[-1] == 2 * len(keys)
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    A = set(zip(rows.tolist(), col.tolist()))
    assert all((c, r) in A for r, c in A)
    for i in range(0, n, 97):
        seg = col[rowptr[i]:rowptr[i + 1]]
        assert np.all(np.diff(seg) > 0)
    assert np.array_equal(np.diff(rowptr), I.degrees(n, keys).astype(int))


def test_generators_deterministic():
    a = I.chung_lu(2000, 2.1, 8.0, 5)
    b = I.chung_lu(2000, 2.1, 8.0, 5)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])
    s1 = I.sbm(500, 5, 0.1, 0.01, 3)
    s2 = I.sbm(500, 5, 0.1, 0.01, 3)
    assert np.array_equal(s1[1], s2[1]) and np.array_equal(s1[2], s2[2])


def test_cap_clusters_reference_rule():
    # alm_vcs.py:1133-1137: keep the max-1 largest-volume clusters, merge the rest into cluster max-1
    lab = np.array([0, 0, 1, 2, 2, 2, 3, 4])
    d = np.array([1, 1, 5, 1, 1, 1, 2, 9], float)
    out = I.cap_clusters(lab, d, 3)
    # volumes: 0->2, 1->5, 2->3, 3->2, 4->9 ; keep 4 (new 0), 1 (new 1); rest -> 2
    assert out.tolist() == [2, 2, 1, 2, 2, 2, 2, 0]


def test_canonical_triangles_merge():
    TI = np.array([[5, 1, 2], [2, 1, 5], [0, 3, 4]])
    T2, lam = I.canonical_triangles(TI, 6, np.array([1.0, 2.0, 0.5]))
    assert T2.tolist() == [[0, 3, 4], [2, 1, 5]]
    assert lam.tolist() == [0.5, 3.0]


def test_levels_rq_dsq():
    lab = np.array([1, 1, 0, 0, 0])
    d = np.array([2.0, 3.0, 1.0, 1.0, 4.0])
    lv = I.pair_level(lab, d)
    assert lv.P == 2
    assert np.allclose(lv.dsq, [1 + 1 + 16, 4 + 9])
    assert np.allclose(lv.rq, [3 / 36.0, 2 / 25.0])


def test_random_ws_contains_own_sorted():
    lab = np.random.default_rng(0).integers(0, 8, 1000).astype(np.int32)
    ptr, col = I.random_ws(lab, 8, 4, 1)
    vl = I.vcs_level(lab, np.ones(1000), ptr, col)
    own = vl.own()
    assert np.all(own >= 0)
    for i in range(0, 1000, 37):
        seg = col[ptr[i]:ptr[i + 1]]
        assert np.all(np.diff(seg) > 0) and lab[i] in seg


def test_wedge_triangles_are_wedges():
    n, keys = I.chung_lu(4000, 2.1, 10.0, 2)
    inst = I.build_instance(n, keys)
    TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, 5000, 3)
    assert len(TI) == 5000 and np.all(TI[:, 0] < TI[:, 2])
    E = set(map(int, keys))
    for a, j, c in TI[:200]:
        assert min(a, j) * n + max(a, j) in E and min(c, j) * n + max(c, j) in E
    kk = (TI[:, 0].astype(np.int64) * n + TI[:, 1]) * n + TI[:, 2]
    assert len(np.unique(kk)) == len(kk)


def test_wedge_triangles_row_path():
    """The unpacked path (taken for n >= 2^21, where (lo n + j) n + hi overflows int64) yields distinct, sorted,
    canonical wedges of the graph, as many as asked."""
    n, keys = I.chung_lu(4000, 2.1, 10.0, 2)
    inst = I.build_instance(n, keys)
    TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, 5000, 3, packed=False)
    assert len(TI) == 5000 and np.all(TI[:, 0] < TI[:, 2])
    E = set(map(int, keys))
    for a, j, c in TI:
        assert min(a, j) * n + max(a, j) in E and min(c, j) * n + max(c, j) in E
    kk = (TI[:, 0].astype(np.int64) * n + TI[:, 1]) * n + TI[:, 2]
    assert np.all(np.diff(kk) > 0)


def test_search_sorted_queries_exact():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 1000, 100000):
        for w in (rng.random(n), (n / (np.arange(n) + 1.0)) ** 0.9,
                  np.where(rng.random(n) < 0.5, 0.0, 1.0) + (np.arange(n) == n - 1)):
            cw = np.cumsum(w)
            cw /= cw[-1]
            q = np.r_[rng.random(20000), 0.0, np.nextafter(1.0, 0.0), cw[:5], 1.0]
            assert np.array_equal(I._search_sorted_queries(cw, q), np.searchsorted(cw, q, side="right"))


def test_label_propagation_dense_equals_sorted():
    n, keys = I.chung_lu(20000, 2.1, 12.0, 4)
    inst = I.build_instance(n, keys)
    for P, it, seed in ((64, 2, 100), (7, 6, 3)):
        a = I.label_propagation(inst.rowptr, inst.col, n, seed, P, it, dense=True)
        b = I.label_propagation(inst.rowptr, inst.col, n, seed, P, it, dense=False)
        assert np.array_equal(a, b)


def test_csr_from_unsorted_weighted_keys():
    n, keys = I.chung_lu(3000, 2.1, 8.0, 5)
    w = np.arange(1, len(keys) + 1, dtype=np.float64)
    perm = np.random.default_rng(1).permutation(len(keys))
    r0, c0, w0 = I.csr_from_keys(n, keys, w)
    r1, c1, w1 = I.csr_from_keys(n, keys[perm], w[perm])
    assert np.array_equal(r0, r1) and np.array_equal(c0, c1) and np.array_equal(w0, w1)
    rows = np.repeat(np.arange(n), np.diff(r0))
    lo, hi = np.minimum(rows, c0), np.maximum(rows, c0)
    kk = lo.astype(np.int64) * n + hi
    assert np.array_equal(w0, w[np.searchsorted(keys, kk)])
