"""Drop-in for the reference's reduction stage, `python _core/reduce_exact.py <tag> [label files ...]` (launched by
Workspace.reduce, pipeline.py:96-99): exact optimum-preserving reduction of data/<tag>/ into data/<tag>__rx/ with the
device kernels of liblbhip (lb_reduce_exact).

Rules (reduce_exact.py:1-8): a degree-1 vertex joins its neighbour; adjacent true twins (N[u] = N[v]) with
d^2 < 2m and d > 1 are merged. The reduced instance is the group graph with integer edge weights, volumes, L = 2m and
C0 = 2 * inside edges, so 2m Q(X) = C0 + <B', X'>. Writes graph.json, keys.npy, weights.npy, vol.npy, group_of.npy,
meta.json and the restricted label files (the label of a group is the label of its first member), checking for every
label file that the original modularity with the rules enforced equals the reduced modularity exactly (Fractions),
as reduce_exact.py:49-62 does. Files, keys and dtypes are the reference's.
"""
from __future__ import annotations

import json
import os
import sys
from fractions import Fraction

import numpy as np

from .engine import Engine


def reference_hash(n):
    """The candidate-filter hash weights of reduce_exact.py:31 (numpy PCG64, seed 11); the result does not depend on
    them, they are passed so the device filters exactly the reference's candidate pairs."""
    return np.random.default_rng(11).integers(1, 2 ** 62, size=n, dtype=np.int64).astype(np.uint64)


def reduce_graph(n, keys, device=0, rhash="reference"):
    """Device reduction of a graph given by sorted unique keys u*n+v (u < v)."""
    eng = Engine(r=8, dtype="f32", device=device)
    try:
        rh = reference_hash(n) if isinstance(rhash, str) and rhash == "reference" else rhash
        return eng.reduce_exact(n, keys, rh)
    finally:
        eng.close()


def restrict_labels(lab, grp, N):
    """Label of each group = label of its first member (reduce_exact.py:53-55)."""
    order = np.argsort(grp, kind="stable")
    gs = grp[order]
    firsts = order[np.r_[True, gs[1:] != gs[:-1]]] if len(gs) else order
    return lab[firsts]


def check_labels(lab, res, n, keys):
    """(Q original, Q with the rules enforced, Q reduced) as Fractions; the last two must be equal
    (reduce_exact.py:55-61)."""
    u, v = keys // n, keys % n
    m = len(keys)
    L = 2 * m
    d = np.bincount(np.r_[u, v], minlength=n)
    grp, N, w, kk = res["group_of"], res["N"], res["weights"], res["keys"]
    lr = restrict_labels(lab, grp, N)
    lab2 = lr[grp]
    P = int(lab2.max()) + 1
    volc = np.bincount(lab2, weights=d, minlength=P).astype(np.int64)
    ein = int((lab2[u] == lab2[v]).sum())
    Qo = Fraction(2 * ein, L) - Fraction(int((volc.astype(object) ** 2).sum()), L * L)
    volr = np.bincount(lr, weights=res["vol"], minlength=P).astype(np.int64)
    ein_r = int(w[lr[kk // N] == lr[kk % N]].sum()) if len(kk) else 0
    Qr = Fraction(2 * res["inside_edges"] + 2 * ein_r, L) - Fraction(int((volr.astype(object) ** 2).sum()), L * L)
    volo = np.bincount(lab, weights=d).astype(np.int64)
    Qorig = Fraction(2 * int((lab[u] == lab[v]).sum()), L) - Fraction(int((volo.astype(object) ** 2).sum()), L * L)
    return Qorig, Qo, Qr, lr


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        raise SystemExit("usage: python -m lagbound_amd.reduce <tag> [label files ...]")
    tag = argv[0]
    D = f"data/{tag}/"
    Dz = f"data/{tag}__rx/"
    os.makedirs(Dz, exist_ok=True)
    n = int(json.load(open(D + "graph.json"))["n"])
    keys = np.load(D + "keys.npy").astype(np.int64)
    m = len(keys)
    L = 2 * m
    res = reduce_graph(n, keys)
    N, kk, w, vol, grp = res["N"], res["keys"], res["weights"], res["vol"], res["group_of"]
    C0e = res["inside_edges"]
    assert int(vol.sum()) == L and int(w.sum()) + C0e == m
    json.dump({"n": int(N), "m": int(len(kk))}, open(Dz + "graph.json", "w"))
    np.save(Dz + "keys.npy", kk.astype(np.int64))
    np.save(Dz + "weights.npy", w.astype(np.float64))
    np.save(Dz + "vol.npy", vol.astype(np.int64))
    np.save(Dz + "group_of.npy", grp.astype(np.int64))
    json.dump({"L_orig": int(L), "leaf_edges": int(C0e), "inside_edges": int(C0e), "n_orig": int(n), "m_orig": int(m),
               "leaf_unions": res["leaf_unions"], "twin_unions": res["twin_unions"]}, open(Dz + "meta.json", "w"))
    print(f"{tag}: n {n} -> {N} (leaf edges {res['leaf_unions']}, twin pairs {res['twin_unions']}), m {m} -> "
          f"{len(kk)} weighted (max w {int(w.max()) if len(w) else 0}), inside edges {C0e}, C0/L {2 * C0e / L:.6f}")
    for f in argv[1:]:
        lab = np.load(f).astype(np.int64)
        Qorig, Qo, Qr, lr = check_labels(lab, res, n, keys)
        assert Qo == Qr, (float(Qo), float(Qr))
        np.save(Dz + os.path.basename(f), lr)
        print(f"  {os.path.basename(f)}: original Q {float(Qorig):.6f}; with rules enforced {float(Qo):.6f} == "
              f"reduced {float(Qr):.6f} OK")
    return 0


if __name__ == "__main__":
    sys.exit(main())
