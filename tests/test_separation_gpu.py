"""Bit-exact parity of the device triangle separation (lb_separate_triangles) with the oracle on the same (stored) V:
identical candidate lists in identical order and identical FP64 violation values, for both families, with binding
histogram caps, hub centres cut to their top-maxh arms, and the three storage types."""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance

pytestmark = pytest.mark.gpu


def _engine(inst, V, dtype, r):
    from lagbound_amd.engine import Engine
    eng = Engine(r=r, dtype=dtype)
    eng.set_instance(inst)
    eng.set_V(V)
    return eng


def _same(got, want):
    TI, val, vmin = got
    TIo, valo, vmino = want
    assert vmin == vmino
    assert TI.shape == TIo.shape, (TI.shape, TIo.shape)
    assert np.array_equal(TI, TIo)
    assert np.array_equal(val, valo)


@pytest.mark.parametrize("dtype", ["f32", "f64", "bf16"])
@pytest.mark.parametrize("fam", [0, 1])
def test_chung_lu_bit_exact(dtype, fam):
    inst, V = chung_lu_instance(n=20000, seed=12, k=16)
    eng = _engine(inst, V, dtype, 16)
    Vs = eng.get_V()
    for cap, thresh, maxh in ((2000, 0.2, 100), (10 ** 6, 0.1, 8), (50, 0.3, 100)):
        print(f"[sep] dtype={dtype} fam={fam} cap={cap} thresh={thresh} maxh={maxh}", flush=True)
        got = eng.separate_triangles(fam, thresh=thresh, maxh=maxh, cap=cap)
        want = O.separate_c(inst, Vs, fam, thresh=thresh, maxh=maxh, cap=cap)
        _same(got, want)
    eng.close()


@pytest.mark.parametrize("fam", [0, 1])
def test_small_uncapped(fam):
    inst, V = small_instance(n=60, seed=4, k=8, p_in=0.4, p_out=0.1)
    eng = _engine(inst, V, "f64", 8)
    Vs = eng.get_V()
    got = eng.separate_triangles(fam, thresh=0.05, maxh=10 ** 6, maxk=2000, hub2=40, cap2=50, cap=10 ** 6, nchunk=7, nb=50)
    want = O.separate_c(inst, Vs, fam, thresh=0.05, maxh=10 ** 6, maxk=2000, hub2=40, cap2=50, cap=10 ** 6, nchunk=7, nb=50)
    _same(got, want)
    eng.close()


def test_nothing_selected():
    """No arm passes the threshold: every tail count is 0 <= cap, so the first bin qualifies (vmin = vtol, as the
    reference's cum <= cap rule gives) and no candidate is returned."""
    inst, V = small_instance(n=40, seed=5, k=8)
    eng = _engine(inst, V, "f32", 8)
    TI, val, vmin = eng.separate_triangles(0, thresh=0.99999, cap=100)
    want = O.separate_c(inst, eng.get_V(), 0, thresh=0.99999, cap=100)
    assert len(TI) == 0 and len(want[0]) == 0 and vmin == want[2] == 1e-3
    eng.close()
