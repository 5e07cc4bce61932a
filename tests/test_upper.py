"""The whole upper-bound chain (reduction -> levels -> main -> proximal -> estimate -> lifted sparse certificate ->
independent verifier) on the reference's own known answers, on the CPU operators (device engine replaced by the FP64
oracle):
Zachary's karate club, optimum 0.4197896120973044 (tests/test_bounds.py:10), Les Miserables, optimum 72259/129032
with the reference's bar U < 0.58 (tests/test_bounds.py:78-82), and the graphs the reductions solve exactly,
K12 -> 0 and K5 + K6 -> 0.48 (tests/test_bounds.py:69-75)."""
import itertools

import networkx as nx
import numpy as np
import pytest

import oracle as O
from lagbound_amd import alm, upper
from lagbound_amd import bound as B
from lagbound_amd import reduce as RD
from oracle_engine import OracleEngine, OracleLagrangian

KARATE_OPT = 0.4197896120973044


def nx_keys(g):
    g = nx.convert_node_labels_to_integers(g)
    n = g.number_of_nodes()
    keys = np.unique([min(a, b) * n + max(a, b) for a, b in g.edges() if a != b]).astype(np.int64)
    return n, keys


@pytest.fixture
def cpu_chain(monkeypatch):
    monkeypatch.setattr(alm, "Engine", OracleEngine)
    monkeypatch.setattr(B, "DeviceLagrangian", lambda inst, r, dtype="f64", device=0: OracleLagrangian(inst, r))
    monkeypatch.setattr(RD, "reduce_graph", lambda n, keys, device=0, rhash="reference": O.reduce_c(n, keys))


@pytest.mark.parametrize("g, q", [(nx.complete_graph(12), 0.0),
                                  (nx.disjoint_union(nx.complete_graph(5), nx.complete_graph(6)), 0.48)])
def test_solved_by_reduction(cpu_chain, g, q):
    res = upper.upper_bound_small(*nx_keys(g))
    assert res["status"] == "solved_by_reduction"
    assert q <= res["value"] <= q + 1e-15


def test_karate_certified_at_the_optimum(cpu_chain):
    res = upper.upper_bound_small(*nx_keys(nx.karate_club_graph()), outer=400, prox_outer=150)
    assert res["status"] == "verified", res
    assert KARATE_OPT <= res["value"] <= KARATE_OPT + 2e-3, res


def test_les_miserables_reduces_to_46_and_certifies_below_058(cpu_chain):
    """tests/test_bounds.py:78-82: Les Miserables reduces to 46 nodes; 72259/129032 <= U < 0.58."""
    res = upper.upper_bound_small(*nx_keys(nx.les_miserables_graph()), outer=400, prox_outer=150)
    assert res["reduced_n"] == 46
    assert res["status"] == "verified", res
    assert 72259 / 129032 <= res["value"] < 0.58, res
