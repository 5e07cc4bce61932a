"""The plain FP64 torch form (the 'vs FP64' comparison baseline) computes the same f and G as the oracle (run on the
CPU here; the bench runs it on the GPU)."""
import numpy as np
import pytest

import oracle as O
from conftest import small_instance, chung_lu_instance


@pytest.mark.parametrize("rho", [0.0, 0.2])
def test_plain_fp64_matches_oracle(rho):
    import torch                     # imported here only: a test process that drives the device library keeps
    from lagbound_amd.baseline_fp64 import PlainFP64   # torch's bundled HIP runtime out (bench runs it apart)
    for inst, V in (small_instance(n=40, seed=1, k=16), chung_lu_instance(n=3000, seed=2, k=16)):
        pl = PlainFP64(inst, device="cpu")
        f, G = pl.eval(torch.as_tensor(V), rho)
        po, Go, _, _, _ = O.eval_c(inst, V, rho)
        assert abs(float(f) - po[0]) <= 1e-10 * max(1.0, abs(po[0]))
        assert np.linalg.norm(G.numpy() - Go) <= 1e-11 * np.linalg.norm(Go)
