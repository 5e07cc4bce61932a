"""Device exact reduction (lb_reduce_exact) bit-exact against the C oracle, and the reduction stage drop-in end to
end (pipeline.reduce -> data/<tag>__rx files identical to the oracle's, label files restricted with the exact Q
check)."""
import itertools
import json
import os

import numpy as np
import pytest

import oracle as O
from lagbound_amd import instances as I
from lagbound_amd import pipeline
from lagbound_amd import reduce as R
from test_reduce import GRAPH, _keys, _random_graph

pytestmark = pytest.mark.gpu


def _same(a, b):
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def _cases():
    yield "fixture", 8, _keys(8, GRAPH)
    yield "K12", 12, _keys(12, list(itertools.combinations(range(12), 2)))
    yield "K5+K6", 11, _keys(11, list(itertools.combinations(range(5), 2)) +
                             list(itertools.combinations(range(5, 11), 2)))
    yield "empty", 5, np.zeros(0, np.int64)
    for s in range(4):
        n, k = _random_graph(7, 0.45, s)
        yield f"random{s}", n, k
    n, k = I.chung_lu(20000, 2.1, 10, 7)
    n2, k2 = I.add_twins_and_leaves(n, k, 400, 600, 8)
    yield "chung-lu-20k", n2, k2


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_device_matches_oracle(case):
    name, n, keys = case
    print(f"[reduce] case {name} n={n} m={len(keys)}", flush=True)
    ref = O.reduce_c(n, keys)
    _same(R.reduce_graph(n, keys), ref)
    _same(R.reduce_graph(n, keys, rhash=None), ref)        # the library's own hash weights


def test_chung_lu_300k_planted():
    n, k = I.chung_lu(300000, 2.1, 13, 11)
    n2, k2 = I.add_twins_and_leaves(n, k, 5000, 20000, 12)
    print(f"[reduce] chung-lu-300k n={n2} m={len(k2)}", flush=True)
    ref = O.reduce_c(n2, k2)
    got = R.reduce_graph(n2, k2)
    _same(got, ref)
    assert got["twin_unions"] > 1000 and got["leaf_unions"] > 20000


def test_reduce_stage_drop_in(tmp_path):
    ws = pipeline.Workspace(str(tmp_path))
    n, k = I.chung_lu(5000, 2.1, 8, 21)
    n2, k2 = I.add_twins_and_leaves(n, k, 100, 200, 22)
    D = os.path.join(ws.data, pipeline.TAG)
    os.makedirs(D)
    json.dump({"n": int(n2), "m": int(len(k2))}, open(os.path.join(D, "graph.json"), "w"))
    np.save(os.path.join(D, "keys.npy"), k2)
    lab = np.random.default_rng(0).integers(0, 12, n2)
    np.save(os.path.join(D, "lab.npy"), lab)
    print("[reduce] stage drop-in", flush=True)
    tag = ws.reduce([os.path.join(D, "lab.npy")])
    Dz = os.path.join(ws.data, tag)
    ref = O.reduce_c(n2, k2)
    assert json.load(open(os.path.join(Dz, "graph.json"))) == {"n": ref["N"], "m": len(ref["keys"])}
    assert np.array_equal(np.load(os.path.join(Dz, "keys.npy")), ref["keys"])
    assert np.array_equal(np.load(os.path.join(Dz, "weights.npy")), ref["weights"])
    assert np.array_equal(np.load(os.path.join(Dz, "vol.npy")), ref["vol"])
    assert np.array_equal(np.load(os.path.join(Dz, "group_of.npy")), ref["group_of"])
    meta = json.load(open(os.path.join(Dz, "meta.json")))
    assert meta["L_orig"] == 2 * len(k2) and meta["leaf_edges"] == ref["inside_edges"]
    assert meta["twin_unions"] == ref["twin_unions"] and meta["leaf_unions"] == ref["leaf_unions"]
    lr = np.load(ws.reduced_labels(os.path.join(D, "lab.npy")))
    assert np.array_equal(lr, R.restrict_labels(lab, ref["group_of"], ref["N"]))
    assert "OK" in open(os.path.join(ws.runs, "reduce.log")).read()
