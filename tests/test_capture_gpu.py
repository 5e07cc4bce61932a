"""Captured evaluation (HIP graph) on the device: every replay equals the eager evaluation at the same inputs, with
eager work and synchronisations between replays; a new V is read by the next replay; a batch context captures its
batch evaluation the same way."""
import numpy as np
import pytest

from lagbound_amd.batch import BatchEngine, Packed
from lagbound_amd.engine import Engine
from batch_case import make_batch
from conftest import chung_lu_instance

pytestmark = pytest.mark.gpu


def test_single_graph_replays_equal_eager():
    inst, V = chung_lu_instance(n=4000, seed=3, k=16)
    e = Engine(r=16, dtype="f32")
    e.set_instance(inst)
    e.set_V(V)
    ref = e.eval(0.02)
    Gref = e.get_G()
    e.capture_eval(0.02)
    for k in range(3):
        print(f"[capture] single replay {k}", flush=True)
        got = e.replay_eval(1)
        assert np.array_equal(got, ref)
        assert np.array_equal(e.get_G(), Gref)
        other = e.eval(0.0)                      # eager work between replays
        assert not np.array_equal(other, ref)
        e.sync()
    V2 = np.roll(V, 1, axis=0)
    e.set_V(V2)
    got2 = e.replay_eval(1)
    G2 = e.get_G()
    eager2 = e.eval(0.02)
    assert np.array_equal(got2, eager2) and np.array_equal(G2, e.get_G())
    e.set_lam(inst.lam * 0.5)                    # a set_* call ends the capture
    with pytest.raises(Exception):
        e.replay_eval(1)
    e.close()


def test_batch_replays_equal_eager():
    p = Packed(make_batch(6, seed=4))
    V = np.random.default_rng(0).standard_normal((p.n, 8))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    be = BatchEngine(p, r=8, dtype="f32")
    be.set_V(V)
    ref = be.eval(0.02)
    Gref = be.get_G()
    be.eng.capture_eval(0.02)
    for k in range(3):
        print(f"[capture] batch replay {k}", flush=True)
        got = be.eng.replay_eval(1)
        # the batch's cluster sums use atomics: equal to rounding, not bit for bit
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-9)
        assert np.allclose(be.get_G(), Gref, rtol=1e-5, atol=1e-6)
        be.eval(0.0)
        be.sync()
    be.close()
