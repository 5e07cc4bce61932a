"""A batch of uneven graphs (SBM and Chung-Lu) with one pair level, multipliers and triangles each, for the batch
tests."""
import numpy as np

from lagbound_amd import instances as I


def make_batch(count=10, seed=0, n_lo=60, n_hi=2500, with_levels=True, tri_per_n=3):
    rng = np.random.default_rng(seed)
    insts = []
    for g in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        if g % 3 == 2:
            n2, keys = I.chung_lu(n, 2.1, 8, seed * 1000 + g)
            lab = I.label_propagation(*I.csr_from_keys(n2, keys, None)[:2], n2, g, P=8, iters=2)
        else:
            B = int(rng.integers(2, 9))
            pin = float(rng.uniform(0.05, 0.3))
            n2, keys, lab = I.sbm(n, B, pin, pin / 25, seed * 1000 + g)
        inst = I.build_instance(n2, keys)
        if with_levels:
            lv = I.pair_level(lab, inst.d)
            S = np.abs(rng.standard_normal((lv.P, lv.P))) * 0.5 / inst.L
            lv.sig = 0.5 * (S + S.T)
            inst.levels = [lv]
        T = tri_per_n * n2
        inst.TI = I.wedge_triangles(inst.rowptr, inst.col, inst.d, T, seed * 1000 + g + 7)
        inst.lam = np.where(rng.random(len(inst.TI)) < 0.5, 0.0, rng.exponential(0.3, len(inst.TI)))
        insts.append(inst)
    return insts
