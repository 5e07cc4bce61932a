"""Pins the oracle's triangle separation (or_separate) against brute-force enumeration on small graphs.

Wedges (a, j, c): every pair of arms a, c of centre j (neighbours with X_aj > thresh), violation
vm = X_aj + X_cj - X_ac - 1 (alm_vcs.py:567-597). One-edge arm (h, j, c): h an arm of j, c a non-neighbour of j
reached through one of the first hub2 arms, X_cj > thresh (alm_vcs.py:608-663). With caps that do not bind, the
selected set is every candidate with vm >= vmin, vmin the histogram threshold, and the histogram counts every
candidate with vm > vtol."""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance


def _dot(a, b):
    s = 0.0
    for x, y in zip(a, b):
        s += x * y
    return s


def _brute_wedges(inst, V, thresh):
    n = inst.n
    out = {}
    for j in range(n):
        nb = inst.col[inst.rowptr[j]:inst.rowptr[j + 1]]
        arms = [(int(i), _dot(V[i], V[j])) for i in nb]
        arms = [(i, x) for i, x in arms if x > thresh]
        for p in range(len(arms)):
            for q in range(p + 1, len(arms)):
                (a, xa), (c, xc) = arms[p], arms[q]
                vm = xa + xc - _dot(V[a], V[c]) - 1.0
                out[(min(a, c), j, max(a, c))] = vm
    return out


def _brute_one_edge(inst, V, thresh):
    n = inst.n
    out = {}
    for j in range(n):
        nbj = set(int(x) for x in inst.col[inst.rowptr[j]:inst.rowptr[j + 1]])
        if not nbj:
            continue
        arms = [(i, _dot(V[i], V[j])) for i in sorted(nbj)]
        arms = [(i, x) for i, x in arms if x > thresh]
        K = {}
        for a, _ in arms:
            for c in inst.col[inst.rowptr[a]:inst.rowptr[a + 1]]:
                c = int(c)
                if c == j or c in nbj:
                    continue
                x = _dot(V[c], V[j])
                if x > thresh:
                    K[c] = x
        for h, xh in arms:
            for c, xc in K.items():
                vm = xh + xc - _dot(V[h], V[c]) - 1.0
                out[(min(h, c), j, max(h, c))] = vm
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("fam", [0, 1])
def test_uncapped_equals_brute_force(seed, fam):
    inst, V = small_instance(n=60, seed=seed, k=4, p_in=0.35, p_out=0.08)
    thresh, vtol = 0.1, 1e-3
    brute = _brute_wedges(inst, V, thresh) if fam == 0 else _brute_one_edge(inst, V, thresh)
    viol = {t: v for t, v in brute.items() if v > vtol}
    TI, val, vmin, hist = O.separate_c(inst, V, fam, thresh=thresh, maxh=10 ** 6, maxk=10 ** 6, hub2=10 ** 6,
                                       cap2=10 ** 6, vtol=vtol, cap=10 ** 6, nchunk=7, nb=50, with_hist=True)
    assert hist.sum() == len(viol)
    got = {tuple(int(x) for x in t): v for t, v in zip(TI, val)}
    want = {t: v for t, v in viol.items() if v >= vmin}
    assert set(got) == set(want)
    for t in got:
        assert got[t] == want[t]


@pytest.mark.parametrize("fam", [0, 1])
def test_cap_selects_the_top_bins(fam):
    inst, V = chung_lu_instance(n=3000, seed=4, k=8)
    TI, val, vmin, hist = O.separate_c(inst, V, fam, thresh=0.2, vtol=1e-3, cap=500, nchunk=64, nb=200,
                                       with_hist=True)
    assert 0 < len(TI) <= 500
    assert np.all(val >= vmin)
    # the histogram tail from vmin's bin holds at most cap, the next lower bin would exceed it
    b0 = int(round((vmin - 1e-3) / (2 - 1e-3) * 200))
    tail = hist.sum(0)[b0:].sum()
    assert tail <= 500 and hist.sum(0)[b0 - 1:].sum() > 500
    assert len(TI) == tail
    assert np.all(TI[:, 0] < TI[:, 2])
    # candidates come chunk by chunk in increasing centre order
    cs = (inst.n + 63) // 64
    assert np.all(np.diff(TI[:, 1] // cs) >= 0)


def test_maxh_cut_keeps_strongest_arms():
    inst, V = chung_lu_instance(n=2000, seed=5, k=8)
    full = O.separate_c(inst, V, 0, thresh=0.0, maxh=10 ** 6, cap=10 ** 7, vtol=1e-3, nchunk=16, nb=100)
    cut = O.separate_c(inst, V, 0, thresh=0.0, maxh=3, cap=10 ** 7, vtol=1e-3, nchunk=16, nb=100)
    fs = set(map(tuple, full[0].tolist()))
    cs_ = set(map(tuple, cut[0].tolist()))
    assert cs_ <= fs and len(cs_) < len(fs)
    # every kept wedge uses two of the three strongest arms of its centre
    for a, j, c in list(cs_)[:200]:
        nb = inst.col[inst.rowptr[j]:inst.rowptr[j + 1]]
        xs = sorted(((_dot(V[i], V[j]), i) for i in nb), reverse=True)[:3]
        top = {i for _, i in xs}
        assert a in top and c in top
