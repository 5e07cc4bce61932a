"""The device solver accepts the reference pipeline's production flag sets and rejects what it does not run."""
import numpy as np
import pytest

from lagbound_amd import alm, pipeline
from lagbound_amd import instances as I


def test_production_flags_parse():
    main = pipeline.solver_args("main", ["a.npy", "b.npy"], "c.npy", 0.5, 100.0, 3000, 512, init="p.npy")
    prox = pipeline.solver_args("proximal", ["a.npy"], "c.npy", 0.5, 100.0, 3000, 512, resume="r.npz",
                                pair_weight=0.01, proximal=True)
    for args in (main, prox):
        args = list(args)
        args[args.index("--k") + 1] = "16"
        a = alm.parse(args)
        assert a.nbatch == 512 and a.est_every == 0 and a.vcs_add == 10000000
    a = alm.parse([x if x != "32" else "16" for x in main])
    assert a.tri_add == 600000 and a.tri1_add == 300000 and a.tri_every == 9 and a.pw_rule == "default"


@pytest.mark.parametrize("bad", [["--mom", "0.5"], ["--est_every", "10"], ["--ne_split", "1"], ["--k", "12"],
                                 ["--nbr_cut", "1"], ["--option", "sweep_fused"], ["--option", "=1"],
                                 ["--option", "sweep_fused=yes"]])
def test_unsupported_flags_rejected(bad):
    with pytest.raises(SystemExit):
        alm.parse(["--tag", "g", "--save", "x.npz", "--est_every", "0", "--k", "16"] + bad)


def test_device_options_parse():
    a = alm.parse(["--tag", "g", "--est_every", "0", "--k", "16", "--option", "sweep_fused=0", "--option", "vcs_grid=-1"])
    assert a.option == ["sweep_fused=0", "vcs_grid=-1"]


def test_load_instance_reduced(tmp_path):
    """Reduced-instance files (weights, vol, meta) load as the reference's A, d, L, C0 (alm_vcs.py:1103-1117)."""
    import json
    D = tmp_path / "g__rx"
    D.mkdir()
    n = 5
    keys = np.array([0 * n + 1, 1 * n + 2, 2 * n + 3, 3 * n + 4, 0 * n + 4], np.int64)
    w = np.array([1.0, 2.0, 1.0, 3.0, 1.0])
    vol = np.array([3, 4, 4, 5, 6], np.int64)
    json.dump({"n": n, "m": 5}, open(D / "graph.json", "w"))
    np.save(D / "keys.npy", keys)
    np.save(D / "weights.npy", w)
    np.save(D / "vol.npy", vol)
    json.dump({"L_orig": int(vol.sum()), "leaf_edges": 2, "inside_edges": 2}, open(D / "meta.json", "w"))
    inst = alm.load_instance(str(D))
    assert inst.n == 5 and inst.L == 22 and inst.C0 == 4.0 and np.array_equal(inst.d, vol)
    A = inst.dense_A()
    assert A[1, 2] == 2.0 and A[2, 1] == 2.0 and A[3, 4] == 3.0 and A.sum() == 2 * w.sum()


def test_init_partition_colouring():
    n_, keys, lab = I.sbm(400, 8, 0.2, 0.01, 1)
    inst = I.build_instance(n_, keys)
    V = alm.init_partition_V(inst, lab, 4, np.random.default_rng(0), lambda *s: None)
    cols = np.argmax(V, axis=1)
    # every cluster sits on one colour, and the noise is small
    for q in range(8):
        assert len(set(cols[lab == q].tolist())) == 1
    assert np.abs(V - np.eye(4)[cols]).max() < 0.01


def test_merge_ws_from_empty_working_set():
    """First VCS separation starts from an empty working set: kept values are gathered only for kept entries."""
    from lagbound_amd.alm import _merge_ws, _take_kept
    from lagbound_amd import instances as I
    lab = np.array([0, 1, 1, 0, 2])
    vl = I.VCSLevelData(lab=lab, P=3, rq=np.ones(3), ptr=np.zeros(6, np.int64), col=np.zeros(0, np.int32),
                        sig=np.zeros(0))
    ptr, col, src = _merge_ws(vl, np.zeros(0, bool), np.array([3, 0, 3]), np.array([2, 1, 0]))
    assert ptr.tolist() == [0, 1, 1, 1, 3, 3]
    assert col.tolist() == [1, 0, 2]
    assert (src == -1).all()
    assert _take_kept(vl.sig, src, np.float64).tolist() == [0.0, 0.0, 0.0]
    # one kept entry among new ones
    vl2 = I.VCSLevelData(lab=lab, P=3, rq=np.ones(3), ptr=np.array([0, 0, 1, 1, 2, 2]), col=np.array([2, 1], np.int32),
                         sig=np.array([0.5, 0.25]))
    ptr, col, src = _merge_ws(vl2, np.array([False, True]), np.array([1]), np.array([0]))
    assert ptr.tolist() == [0, 0, 1, 1, 2, 2]
    assert col.tolist() == [0, 1]
    assert _take_kept(vl2.sig, src, np.float64).tolist() == [0.0, 0.25]


def test_merge_ws_matches_full_sort():
    """The sorted merge equals a stable sort of [kept, new] by (i, R) on random working sets, including new entries
    that repeat a kept one (kept first) and repeats among the new ones (their given order)."""
    from lagbound_amd.alm import _merge_ws
    from lagbound_amd import instances as I
    rng = np.random.default_rng(3)
    for trial in range(20):
        n, P = int(rng.integers(1, 60)), int(rng.integers(1, 9))
        cnt = rng.integers(0, P + 1, n)
        cols = [rng.choice(P, size=k, replace=False) for k in cnt]
        if trial % 2 == 0:                      # sorted rows (as the solver writes them); odd trials: unsorted
            cols = [np.sort(x) for x in cols]
        ptr0 = np.zeros(n + 1, np.int64)
        np.cumsum(cnt, out=ptr0[1:])
        col0 = np.concatenate(cols).astype(np.int32) if cnt.sum() else np.zeros(0, np.int32)
        vl = I.VCSLevelData(lab=rng.integers(0, P, n), P=P, rq=np.ones(P), ptr=ptr0, col=col0, sig=np.zeros(len(col0)))
        keep = rng.random(len(col0)) < 0.6
        m = int(rng.integers(0, 3 * n))
        ni, nR = rng.integers(0, n, m), rng.integers(0, P, m)
        ptr, col, src = _merge_ws(vl, keep, ni, nR)
        rows_old = np.repeat(np.arange(n), np.diff(ptr0))
        ko = np.flatnonzero(keep)
        r = np.r_[rows_old[ko], ni]
        c = np.r_[col0[ko].astype(np.int64), nR]
        s_ = np.r_[ko, np.full(m, -1)]
        o = np.argsort(r * P + c, kind="stable")
        want_ptr = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(r, minlength=n), out=want_ptr[1:])
        assert np.array_equal(ptr, want_ptr) and np.array_equal(col, c[o]) and np.array_equal(src, s_[o]), trial


def test_new_triangles_matches_filter_then_unique():
    """_new_triangles equals dropping pool members first and deduplicating after (the reference's order of steps)
    on random candidates with repeats and pool hits."""
    from lagbound_amd.alm import _new_triangles
    rng = np.random.default_rng(1)
    for t in range(100):
        n, m, T = int(rng.integers(3, 40)), int(rng.integers(0, 60)), int(rng.integers(0, 50))
        new, val, TI = rng.integers(0, n, (m, 3)), rng.random(m), rng.integers(0, n, (T, 3))
        if m and T:
            TI[:min(T, m) // 2] = new[:min(T, m) // 2]
        got_new, got_val = _new_triangles(new, val, TI, n)
        kn = (new[:, 0] * n + new[:, 1]) * n + new[:, 2]
        pool = set(((TI[:, 0] * n + TI[:, 1]) * n + TI[:, 2]).tolist())
        keep = np.array([k not in pool for k in kn.tolist()], bool)
        k2, v2, n2 = kn[keep], val[keep], new[keep]
        _, f = np.unique(k2, return_index=True)
        assert np.array_equal(got_new, n2[f]) and np.array_equal(got_val, v2[f]), t


def test_new_triangles_keeps_the_int64_wraparound():
    """Above n = 2^21 the packed key (a n + b) n + c overflows int64; the pool filter keeps numpy's wrapped keys (the
    reference's key, alm_vcs.py:1751-1752), so its set and order equal a numpy restatement with the same keys."""
    from lagbound_amd.alm import _new_triangles
    rng = np.random.default_rng(5)
    n = 3_000_000
    new = rng.integers(0, n, (500, 3))
    new[100:150] = new[:50]                              # repeats
    TI = rng.integers(0, n, (300, 3))
    TI[:40] = new[200:240]                              # pool hits
    val = rng.random(500)
    got_new, got_val = _new_triangles(new, val, TI, n)
    with np.errstate(over="ignore"):
        kn = (new[:, 0] * n + new[:, 1]) * n + new[:, 2]
        kp = (TI[:, 0] * n + TI[:, 1]) * n + TI[:, 2]
    uk, fi = np.unique(kn, return_index=True)
    keep = ~np.isin(uk, kp)
    assert np.array_equal(got_new, new[fi[keep]]) and np.array_equal(got_val, val[fi[keep]])


def test_host_argsorts_equal_numpy_stable():
    """liblbhost's stable argsorts (counting sort of bounded keys, radix sort of int64 keys) return numpy's
    argsort(kind="stable") permutation, ties in index order, including empty input, one key, the int64 extremes and
    heavy duplication; a key outside [0, kmax) is an error."""
    from lagbound_amd import _host as H
    rng = np.random.default_rng(5)
    for m, kmax in ((0, 5), (1, 1), (1000, 7), (50000, 40000)):
        k = rng.integers(0, kmax, m)
        assert np.array_equal(H.argsort_bounded(k, kmax), np.argsort(k, kind="stable"))
    for k in (np.zeros(0, np.int64), rng.integers(-2 ** 62, 2 ** 62, 20000), rng.integers(0, 10, 20000),
              np.array([5, -3, 5, 2 ** 63 - 1, -2 ** 63, 0], np.int64)):
        assert np.array_equal(H.argsort_i64(k), np.argsort(k, kind="stable"))
    with pytest.raises(ValueError):
        H.argsort_bounded(np.array([3]), 3)
    with pytest.raises(ValueError):
        H.argsort_bounded(np.array([-1]), 3)


def test_tie_order_equals_lexsort():
    """Candidates in descending score order with runs of tied scores: liblbhost's run-local sort gives
    np.lexsort((R, i, -sc)); an order that is not descending falls back to the full lexsort."""
    from lagbound_amd import _host as H
    rng = np.random.default_rng(8)
    for m in (0, 1, 2, 5000):
        sc = -np.sort(-rng.integers(0, 40, m).astype(np.float64))       # many ties
        i = rng.integers(0, 300, m)
        R = rng.integers(0, 64, m)
        assert np.array_equal(H.tie_order(sc, i, R), np.lexsort((R, i, -sc)))
    sc = np.array([3.0, 1.0, 2.0, 2.0])
    i = np.array([4, 3, 2, 1])
    R = np.zeros(4, np.int64)
    assert np.array_equal(H.tie_order(sc, i, R), np.lexsort((R, i, -sc)))


def test_row_gather_equals_numpy():
    """liblbhost's row gather (rows of a CSR taken in a permutation order) against its numpy statement, with empty
    rows, a non-zero first pointer and no rows at all; a bad order is an error."""
    from lagbound_amd import _host as H
    rng = np.random.default_rng(12)
    for n, base in ((0, 0), (1, 3), (500, 7)):
        cnt = rng.integers(0, 6, n)
        ptr = np.full(n + 1, base, np.int64)
        np.cumsum(cnt, out=ptr[1:])
        ptr[1:] += base
        order = rng.permutation(n)
        c2 = np.diff(ptr)[order]
        nptr = np.zeros(n + 1, np.int64)
        np.cumsum(c2, out=nptr[1:])
        idx = np.repeat(ptr[order] - nptr[:-1], c2) + np.arange(nptr[-1], dtype=np.int64)
        a, b = H.row_gather(ptr, order)
        assert np.array_equal(a, nptr) and np.array_equal(b, idx)
    with pytest.raises(ValueError):
        H.row_gather(np.array([0, 1, 2], np.int64), np.array([0, 2], np.int64))
    with pytest.raises(ValueError):                      # a repeated row: more entries than the CSR holds
        H.row_gather(np.array([0, 1, 3], np.int64), np.array([1, 1], np.int64))


def test_coo_sum_equals_scipy():
    """liblbhost's COO -> CSR conversion with duplicates combined against scipy's (sum_duplicates + sort_indices):
    the same pattern, sums equal up to the order of the adds, absolute sums and counts, and pattern-only input;
    an index out of range is an error."""
    import scipy.sparse as sp
    from lagbound_amd import _host as H
    rng = np.random.default_rng(14)
    for N, m in ((1, 0), (7, 40), (300, 5000)):
        r = rng.integers(0, N, m)
        c = rng.integers(0, N, m)
        v = rng.standard_normal(m)
        ip, ix, s, a, n = H.coo_sum(N, r, c, v, abs_sums=True, counts=True)
        for data, got in ((v, s), (np.abs(v), a), (np.ones(m), n)):
            M = sp.csr_matrix((data, (r, c)), shape=(N, N))
            M.sum_duplicates()
            M.sort_indices()
            assert np.array_equal(ip, M.indptr) and np.array_equal(ix, M.indices)
            assert np.allclose(got, M.data, rtol=1e-13, atol=1e-13)
        ip2, ix2, none, _, _ = H.coo_sum(N, r, c, sums=False)
        assert none is None and np.array_equal(ip2, ip) and np.array_equal(ix2, ix)
    with pytest.raises(ValueError):
        H.coo_sum(3, np.array([0, 3]), np.array([0, 1]))


def test_sym_from_lower_equals_scipy():
    """liblbhost's full symmetric CSR of a lower triangle (bucket transpose, no sort) against scipy's construction of
    the same matrix; an entry above the diagonal or unsorted columns are errors."""
    import scipy.sparse as sp
    from lagbound_amd import _host as H
    for N, dens in ((1, 1.0), (40, 0.2), (500, 0.02)):
        A = sp.random(N, N, density=dens, random_state=2, format="csr")
        Lo = sp.tril(A + A.T + sp.eye(N), format="csr")
        Lo.sort_indices()
        ip, ix, v = H.sym_from_lower(N, Lo.indptr, Lo.indices, Lo.data)
        F = (Lo + sp.tril(Lo, -1).T).tocsr()
        F.sort_indices()
        assert np.array_equal(ip, F.indptr) and np.array_equal(ix, F.indices) and np.array_equal(v, F.data)
    with pytest.raises(ValueError):
        H.sym_from_lower(2, np.array([0, 1, 2]), np.array([1, 1]), np.ones(2))      # (0, 1) is above the diagonal
    with pytest.raises(ValueError):
        H.sym_from_lower(2, np.array([0, 0, 2]), np.array([1, 0]), np.ones(2))      # row 1 columns not ascending
