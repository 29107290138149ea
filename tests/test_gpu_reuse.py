"""ReuseMCConfig for a batch of chains (cthyb_solver.py:24-86, 328-332 keep one
configuration per rank between DMFT iterations): every chain's configuration is
saved after one solve and set into the next handle."""

import numpy as np
import pytest

from w2dynamics_amd import problems as P
from w2dynamics_amd.ctqmc import CtqmcGpu

pytestmark = pytest.mark.gpu


def test_batch_configurations_round_trip(tmp_path):
    """Configurations saved from 64 chains and set into a new handle of the
    same model come back unchanged, with the trace, determinant and sign that
    set_config recomputes from scratch equal to the ones the first handle
    carried through its fast updates; the next solve runs from them."""
    prob = P.kanamori_2orb(beta=20.0, TaudiffMax=4.0)
    nch = 64
    print("batch reuse", flush=True)
    a = CtqmcGpu.from_problem(prob, nchains=nch, seed=11)
    a.set_taudiffmax(4.0)
    a.warmup(2000)
    path = str(tmp_path / "configs.npz")
    a.save_configs(path)
    before = [a.get_config(c) for c in range(nch)]
    a.close()
    b = CtqmcGpu.from_problem(prob, nchains=nch, seed=12)
    b.set_taudiffmax(4.0)
    weights = b.load_configs(path)
    assert len(weights) == nch
    assert sum(o["n"] for o in before) > 0
    for c in range(nch):
        o, g, w = before[c], b.get_config(c), weights[c]
        assert g["n"] == o["n"] and g["outer_sst"] == o["outer_sst"], c
        for k in ("taus", "orbs", "spins", "cas"):
            assert np.array_equal(g[k], o[k]), (c, k)
        for k in ("trace_log", "det_log"):
            assert abs(w[k] - o[k]) <= 1e-8 * max(1.0, abs(o[k])), (c, k, w[k], o[k])
        for k in ("trace_sign", "det_sign", "cfgsign"):
            assert w[k] == o[k], (c, k)
    b.warmup(500)
    assert np.all(b.chain_status()["status"] == 0)
    b.close()
