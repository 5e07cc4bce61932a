"""Block-diagonal packing of a batch (host logic): every graph's CSR rows, volumes, level and triangles come back
unchanged from its slice of the packed arrays."""
import numpy as np

from lagbound_amd.batch import Packed
from batch_case import make_batch


def test_packing_round_trip():
    insts = make_batch(7, seed=3)
    p = Packed(insts)
    assert p.G == 7 and p.n == sum(x.n for x in insts)
    so = 0
    for g, x in enumerate(insts):
        o = p.row_off[g]
        sl = p.rows(g)
        rp = p.rowptr[sl.start:sl.stop + 1] - p.rowptr[sl.start]
        assert np.array_equal(rp, x.rowptr)
        cols = p.col[p.rowptr[sl.start]:p.rowptr[sl.stop]] - o
        assert np.array_equal(cols, x.col)
        assert np.array_equal(p.d[sl], x.d)
        assert np.array_equal(p.lab[sl], x.levels[0].lab)
        P = x.levels[0].P
        assert np.array_equal(p.sig[so:so + P * P], x.levels[0].sig.ravel())
        so += P * P
        t0, t1 = p.tri_off[g], p.tri_off[g + 1]
        assert np.array_equal(p.TI[t0:t1] - o, x.TI)
        assert np.array_equal(p.lam[t0:t1], x.lam)
    assert p.L.tolist() == [x.L for x in insts]
