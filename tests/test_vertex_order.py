"""The degree-ordered device copy of an instance (lagbound_amd.vertex_order): the relabelled CSR is P A P^T with
sorted rows, and a `Relabeled` handle over the FP64 oracle engine returns, in the caller's numbering, what the
oracle engine returns on the instance as given -- evaluation parts, G, effective multipliers, VCS state after a PHR
update, sweeps, working-set replacement, both separations."""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance
from oracle_engine import OracleEngine
from lagbound_amd import instances as I
from lagbound_amd.vertex_order import Relabeled, VertexOrder


def _dense(inst):
    A = np.zeros((inst.n, inst.n))
    rows = np.repeat(np.arange(inst.n), np.diff(inst.rowptr))
    A[rows, inst.col] = 1.0 if inst.w is None else inst.w
    return A


@pytest.mark.parametrize("weighted", [False, True])
def test_graph_is_symmetric_permutation(weighted):
    inst, _ = small_instance(n=70, seed=2, weighted=weighted)
    vo = VertexOrder(inst.rowptr)
    deg = np.diff(inst.rowptr)
    assert np.all(np.diff(deg[vo.order]) <= 0)
    assert np.array_equal(vo.order[vo.new_of], np.arange(inst.n))
    rowptr, col, w, d = vo.graph(inst)
    J = I.Instance(n=inst.n, rowptr=rowptr, col=col, w=w, d=d, L=inst.L)
    for i in range(inst.n):
        assert np.all(np.diff(col[rowptr[i]:rowptr[i + 1]]) > 0)
    A = _dense(inst)
    assert np.array_equal(_dense(J), A[np.ix_(vo.order, vo.order)])
    assert np.array_equal(d, inst.d[vo.order])


def _pair(inst, k, order):
    ref = OracleEngine(r=k)
    rel = Relabeled(OracleEngine(r=k), lambda: OracleEngine(r=k), order)
    return ref, rel


@pytest.mark.parametrize("weighted", [False, True])
def test_relabeled_oracle_engine_matches(weighted):
    inst, V = small_instance(n=80, seed=5, k=8, weighted=weighted, ntri=60, vcs_per_row=3)
    ref, rel = _pair(inst, 8, "degree")
    for e in (ref, rel):
        e.set_instance(inst)
        e.set_V(V)
    assert rel.vo is not None and not np.array_equal(rel.vo.order, np.arange(inst.n))
    # parts within 1e-10: the VCS term is a sum of (p^2 - s^2) / (2 rho r) with tiny rho r, whose rounding moves
    # with the row order of the sum
    for rho in (0.0, 0.3):
        p0 = ref.eval(rho)
        p1 = rel.eval(rho)
        assert np.allclose(p1, p0, rtol=1e-10, atol=1e-12)
        assert np.abs(rel.get_G() - ref.get_G()).max() <= 1e-12 * np.abs(ref.get_G()).max()
        for a, b in zip(rel.get_multipliers(), ref.get_multipliers()):
            assert np.allclose(a, b, rtol=1e-12, atol=1e-14)
    # PHR update: VCS sig and ages in the caller's entry order
    for e in (ref, rel):
        e.phr_update()
    s0, a0 = ref.get_vcs_state()
    s1, a1 = rel.get_vcs_state()
    assert np.allclose(s1, s0, rtol=1e-12, atol=1e-14) and np.array_equal(a1, a0)
    # sweeps: orders in the caller's numbering
    orders = np.stack([np.random.default_rng(s).permutation(inst.n) for s in (1, 2)])
    for e in (ref, rel):
        e.sweeps(orders, 16, 0.3)
    assert np.abs(rel.get_V() - ref.get_V()).max() <= 1e-10
    # triangle separation on the caller's graph and V: identical lists (both on the same V; the sweeps above
    # leave the two V within rounding of each other)
    Vc = ref.get_V()
    rel.set_V(Vc)
    for fam in (0, 1):
        a = ref.separate_triangles(fam, thresh=0.1, maxh=6, cap=40)
        b = rel.separate_triangles(fam, thresh=0.1, maxh=6, cap=40)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    # VCS separation: same set, same order (score desc, then (i, R))
    i0, R0, c0, _ = ref.separate_vcs(0, nadd=10 ** 6)
    i1, R1, c1, _ = rel.separate_vcs(0, nadd=10 ** 6)
    assert set(zip(i0.tolist(), R0.tolist())) == set(zip(i1.tolist(), R1.tolist()))
    assert len(c1) == 0 or np.all(np.diff(c1) <= 1e-15 * np.abs(c1).max())
    # working-set replacement and ages, then a new triangle pool
    vl = inst.vcs[0]
    ptr, col, _ = O.vcs_update_ws(vl, np.arange(vl.nnz) % 2 == 0, i0[:7], R0[:7])
    sig = np.linspace(0.0, 1.0, len(col))
    inst2 = I.Instance(n=inst.n, rowptr=inst.rowptr, col=inst.col, w=inst.w, d=inst.d, L=inst.L, C0=inst.C0,
                       levels=inst.levels)
    inst2.vcs = [I.vcs_level(vl.lab, inst.d, ptr, col, sig)]
    age = (np.arange(len(col)) % 5).astype(np.uint8)
    TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, 40, 9)
    lam = np.linspace(0.0, 2.0, len(TI))
    for e in (ref, rel):
        e.set_vcs(inst2)
        e.set_vcs_age(age)
        e.set_triangles(TI, lam)
        e.set_V(V)
    assert np.array_equal(rel.get_vcs_state()[1], age)
    p0, p1 = ref.eval(0.2), rel.eval(0.2)
    assert np.allclose(p1, p0, rtol=1e-10, atol=1e-12)
    for a, b in zip(rel.get_multipliers(), ref.get_multipliers()):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-14)


def test_given_order_is_identity():
    inst, V = small_instance(n=40, seed=1, k=8)
    ref, rel = _pair(inst, 8, "given")
    for e in (ref, rel):
        e.set_instance(inst)
        e.set_V(V)
    assert rel.vo is None
    assert np.array_equal(rel.eval(0.1), ref.eval(0.1))
    assert np.array_equal(rel.get_G(), ref.get_G())
    with pytest.raises(ValueError):
        Relabeled(OracleEngine(r=8), None, "random")


def test_chung_lu_relabel_eval():
    inst, V = chung_lu_instance(n=5000, seed=2, k=16)
    ref, rel = _pair(inst, 16, "degree")
    for e in (ref, rel):
        e.set_instance(inst)
        e.set_V(V)
    p0, p1 = ref.eval(0.02), rel.eval(0.02)
    assert abs(p1[0] - p0[0]) <= 1e-11 * abs(p0[0])
    assert np.abs(rel.get_G() - ref.get_G()).max() <= 1e-11 * np.abs(ref.get_G()).max()
    i0, R0, c0, s0 = ref.separate_vcs(0, nadd=300)
    i1, R1, c1, s1 = rel.separate_vcs(0, nadd=300)
    assert len(i0) == 300 and s0["nviol"] == s1["nviol"]
    assert set(zip(i0.tolist(), R0.tolist())) == set(zip(i1.tolist(), R1.tolist()))
    o = np.lexsort((R0, i0, -c0))
    assert np.array_equal(i1, i0[o]) and np.array_equal(R1, R0[o])
    assert np.allclose(c1, c0[o], rtol=1e-12)
