"""Exact reduction (reduce_exact.py:12-62): the C oracle against an independent restatement (connected components of
the eligible edges, sets instead of hashes), against the reference's own fixtures and known answers, the hash
independence of the result, and the exact modularity identity of the reduced instance (reduce_exact.py:55-61)."""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from lagbound_amd import instances as I
from lagbound_amd import reduce as R


def _keys(n, edges):
    return np.unique([min(a, b) * n + max(a, b) for a, b in edges]).astype(np.int64)


def _random_graph(n, p, seed):
    """The reference's test generator (tests/conftest.py:23-29): G(n, p) plus a leaf and a twin pair."""
    rng = np.random.default_rng(seed)
    edges = [(i, j) for i, j in itertools.combinations(range(n), 2) if rng.random() < p]
    edges += [(0, n), (1, n + 1), (2, n + 1), (1, 2)]
    return n + 2, _keys(n + 2, edges)


def _same(a, b):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


# the reference's fixture (tests/test_verifier.py:40-47): K4 on 0..3 with leaf 7 at 0 and a triangle 4-5-6 off 3
GRAPH = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4), (4, 5), (5, 6), (4, 6), (0, 7)]


def test_reference_fixture():
    r = O.reduce_c(8, _keys(8, GRAPH))
    assert r["group_of"].tolist() == [0, 1, 1, 2, 3, 4, 4, 0]
    assert r["leaf_unions"] == 1 and r["twin_unions"] == 2
    _same(r, O.reduce_numpy(8, _keys(8, GRAPH)))


def test_large_degree_twins_not_merged():
    # tests/test_verifier.py:67-72: 0 and 1 are adjacent twins of degree 4 with 16 >= 2m = 14
    edges = [(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4)]
    r = O.reduce_c(5, _keys(5, edges))
    assert r["twin_unions"] == 0 and r["N"] == 5


def test_non_adjacent_twins_not_merged():
    # tests/test_verifier.py:61-64: 0 and 2 share all neighbours but are not adjacent
    r = O.reduce_c(4, _keys(4, [(0, 1), (1, 2), (2, 3), (3, 0)]))
    assert r["N"] == 4 and r["twin_unions"] == 0 and r["leaf_unions"] == 0


@pytest.mark.parametrize("name,n,edges,N,q", [
    ("K12", 12, list(itertools.combinations(range(12), 2)), 1, Fraction(0)),
    ("K5+K6", 11, list(itertools.combinations(range(5), 2)) + list(itertools.combinations(range(5, 11), 2)), 2,
     Fraction(48, 100)),
])
def test_solved_by_reduction(name, n, edges, N, q):
    """tests/test_bounds.py:69-75: twin merges leave no edge between groups; the optimum is C0/L - sum vol^2/L^2."""
    keys = _keys(n, edges)
    r = O.reduce_c(n, keys)
    assert r["N"] == N and len(r["keys"]) == 0
    L = 2 * len(keys)
    Q = Fraction(2 * r["inside_edges"], L) - Fraction(int((r["vol"].astype(object) ** 2).sum()), L * L)
    assert Q == q


@pytest.mark.parametrize("seed", range(6))
def test_oracle_vs_components_small(seed):
    n, keys = _random_graph(7, 0.45, seed)
    _same(O.reduce_c(n, keys), O.reduce_numpy(n, keys))


@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_vs_components_planted(seed):
    n, k = I.chung_lu(3000, 2.1, 8, seed)
    n2, k2 = I.add_twins_and_leaves(n, k, 60, 80, seed + 10)
    a = O.reduce_c(n2, k2)
    assert a["twin_unions"] > 50 and a["leaf_unions"] > 80
    _same(a, O.reduce_numpy(n2, k2))


def test_hash_does_not_change_result():
    n, k = I.chung_lu(2000, 2.1, 8, 4)
    n2, k2 = I.add_twins_and_leaves(n, k, 40, 40, 5)
    a = O.reduce_c(n2, k2)
    b = O.reduce_c(n2, k2, rhash=(np.arange(n2, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) | np.uint64(1))
    _same(a, b)


def test_empty_and_isolated():
    r = O.reduce_c(5, np.zeros(0, np.int64))
    assert r["N"] == 5 and r["group_of"].tolist() == [0, 1, 2, 3, 4] and len(r["keys"]) == 0
    r = O.reduce_c(4, _keys(4, [(1, 2)]))            # two leaves: one group, isolated 0 and 3 stay
    assert r["group_of"].tolist() == [0, 1, 1, 2] and r["inside_edges"] == 1 and r["vol"].tolist() == [0, 2, 0]


@pytest.mark.parametrize("seed", [0, 3])
def test_reduced_modularity_identity(seed):
    """For any labels: original Q with the rules enforced == reduced Q, exactly (reduce_exact.py:55-61)."""
    n, k = I.chung_lu(1500, 2.1, 8, seed)
    n2, k2 = I.add_twins_and_leaves(n, k, 30, 30, seed + 1)
    res = O.reduce_c(n2, k2)
    lab = np.random.default_rng(seed).integers(0, 7, n2)
    Qorig, Qo, Qr, lr = R.check_labels(lab, res, n2, k2)
    assert Qo == Qr and len(lr) == res["N"]
    # a planted mistake (one cross edge counted inside) breaks the identity
    bad = dict(res, inside_edges=res["inside_edges"] + 1)
    assert R.check_labels(lab, bad, n2, k2)[1] != R.check_labels(lab, bad, n2, k2)[2]
