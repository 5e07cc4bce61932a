import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblbhip.so")


def small_instance(n=40, seed=0, k=8, blocks=4, p_in=0.3, p_out=0.05, ntri=30, weighted=False, sig_scale=0.01,
                   vcs_per_row=3, levels=2):
    """Small SBM instance with two pair levels, one VCS level and a wedge triangle pool."""
    from lagbound_amd import instances as I
    n_, keys, lab = I.sbm(n, blocks, p_in, p_out, seed)
    rng = np.random.default_rng(seed + 17)
    weights = rng.integers(1, 4, len(keys)).astype(float) if weighted else None
    inst = I.build_instance(n_, keys, weights=weights)
    if weighted:
        inst.C0 = 6.0
        inst.L = float(inst.d.sum())
    lab = lab[:n_]
    S = np.abs(rng.standard_normal((blocks, blocks)))
    inst.levels.append(I.pair_level(lab, inst.d, 0.5 * (S + S.T) * sig_scale))
    if levels > 1:
        lab2 = rng.integers(0, 3, n_)
        lab2[:3] = [0, 1, 2]
        S2 = np.abs(rng.standard_normal((3, 3)))
        inst.levels.append(I.pair_level(lab2, inst.d, 0.5 * (S2 + S2.T) * sig_scale))
    if vcs_per_row:
        ptr, col = I.random_ws(lab, blocks, vcs_per_row, seed)
        inst.vcs.append(I.vcs_level(lab, inst.d, ptr, col, rng.exponential(10 * sig_scale, len(col))))
    if ntri:
        TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, ntri, seed)
        inst.TI = TI
        inst.lam = np.where(rng.random(len(TI)) < 0.4, 0.0, rng.exponential(1.0, len(TI)))
    V = I.init_V(n_, k, rng)
    return inst, V


def chung_lu_instance(n=20000, seed=3, k=16, P=16, ws=4, T_per_n=3.0, sig_scale=None, kbar=13.0, gamma=2.1):
    """Power-law instance with hubs (segment path), balanced LPA pair level and VCS level, wedge triangles."""
    from lagbound_amd import instances as I
    n_, keys = I.chung_lu(n, gamma, kbar, seed)
    inst = I.build_instance(n_, keys)
    lab = I.cap_clusters(I.label_propagation(inst.rowptr, inst.col, n_, seed=seed + 1, P=P, iters=2), inst.d, P)
    rng = np.random.default_rng(seed + 5)
    Pq = int(lab.max()) + 1
    scale = (Pq * Pq / inst.L) if sig_scale is None else sig_scale
    S = np.abs(rng.standard_normal((Pq, Pq)))
    inst.levels.append(I.pair_level(lab, inst.d, 0.5 * (S + S.T) * scale))
    ptr, col = I.random_ws(lab, Pq, ws, seed + 7)
    inst.vcs.append(I.vcs_level(lab, inst.d, ptr, col, rng.exponential(1.0, len(col)) * scale))
    TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, int(T_per_n * n_), seed + 9)
    inst.TI = TI
    inst.lam = np.where(rng.random(len(TI)) < 0.5, 0.0, rng.exponential(1.0, len(TI)))
    V = I.init_V(n_, k, rng)
    return inst, V


@pytest.fixture
def small():
    return small_instance()
