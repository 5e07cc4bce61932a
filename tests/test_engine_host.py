"""Host bookkeeping of the device context handle (DeviceEngine) without a device (the library replaced by a stub that accepts every call): the
getters size their output buffers by what the engine uploaded. Regression test: a batch context uploads its
triangles with set_triangles after set_instance(union without triangles); get_multipliers used to size pt by the
instance's triangle count (0) while the library writes one value per uploaded triangle, overrunning the host
buffer."""
import numpy as np

from lagbound_amd import engine as E
from lagbound_amd import instances as I


class _StubLib:
    def __getattr__(self, name):
        return lambda *a: 0


def _engine(r=8):
    e = object.__new__(E.DeviceEngine)
    e.lib_s = 0.0
    e.n_pp = e.n_t = e.n_pv = 0
    e.r = r
    e.dtype = "f32"
    e.ctx = None
    e.L = _StubLib()
    e.inst = None
    return e


def test_getters_sized_by_uploads():
    n, keys = I.chung_lu(300, 2.1, 6, 1)
    union = I.build_instance(n, keys)
    e = _engine()
    e.set_instance(union)
    assert e.n_t == 0
    TI = I.wedge_triangles(union.rowptr, union.col, union.d, 100, 2)
    e.set_triangles(TI, np.zeros(len(TI)))
    pp, pt, pv = e.get_multipliers()
    assert len(pt) == len(TI) == e.n_t and len(pp) == 0 and len(pv) == 0
    lab = np.arange(n) % 3
    union.levels = [I.pair_level(lab, union.d)]
    e.set_levels(union)
    assert e.n_pp == 9 and len(e.get_multipliers()[0]) == 9
