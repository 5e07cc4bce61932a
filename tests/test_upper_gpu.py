"""The whole upper-bound chain on the device (reduction, main and proximal solver phases, estimate) with the lifted
sparse FP64 certificate and the independent verifier, on the reference's known answers: karate club optimum
0.4197896120973044 (tests/test_bounds.py:10) and Les Miserables, 46 reduced nodes and 72259/129032 <= U < 0.58
(tests/test_bounds.py:78-82)."""
import networkx as nx
import pytest

from lagbound_amd import upper
from test_upper import KARATE_OPT, nx_keys

pytestmark = pytest.mark.gpu


def test_karate_device_chain():
    print("[upper] karate", flush=True)
    res = upper.upper_bound_small(*nx_keys(nx.karate_club_graph()), outer=400, prox_outer=150)
    print(f"[upper] karate {res}", flush=True)
    assert res["status"] == "verified"
    assert KARATE_OPT <= res["value"] <= KARATE_OPT + 2e-3


def test_les_miserables_device_chain():
    print("[upper] les miserables", flush=True)
    res = upper.upper_bound_small(*nx_keys(nx.les_miserables_graph()), outer=400, prox_outer=150)
    print(f"[upper] les miserables {res}", flush=True)
    assert res["reduced_n"] == 46 and res["status"] == "verified"
    assert 72259 / 129032 <= res["value"] < 0.58
