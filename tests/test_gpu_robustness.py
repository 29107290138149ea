"""Batch behaviour on the GPU: chains that outgrow their memory, chains that
stop on an error, and the abort word.

A chain that runs out of room stops before its move's first random draw; the
library doubles every chain's capacity and the chain continues, so its decision
sequence is the one of an unlimited run (the oracle's).  With a fixed capacity
the chain stops with W2D_ERR_CAPACITY and the others carry on.
"""

import warnings

import numpy as np
import pytest

from oracle import OracleChain
from test_gpu_parity import _compare
from w2dynamics_amd import problems as P
from w2dynamics_amd.ctqmc import CtqmcGpu

pytestmark = pytest.mark.gpu


def test_capacity_regrowth_keeps_the_decision_sequence():
    prob = P.kanamori_2orb(beta=20.0, TaudiffMax=4.0)
    seed, n = 12345, 6000
    print("regrowth from capacity 2", flush=True)
    gpu = CtqmcGpu.from_problem(prob, nchains=2, seed=seed, capacity=2)
    gpu.set_taudiffmax(4.0)
    rec = gpu.moves_log(0, n)
    ora = OracleChain.from_problem(prob, seed)
    ora.set_taudiffmax(4.0)
    _compare(rec, ora.moves(n), "regrowth")
    st = gpu.chain_status()
    assert st["capacity"] > 2 and st["kfmax"][0].max() <= st["capacity"]
    assert st["kfmax"][0].max() > 2
    for f in range(4):
        a, b = gpu.get_minv(0, f), ora.minv(f)
        if a.size:
            assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(b).max())
    gpu.close()


def test_batched_run_with_regrowth_matches_large_capacity():
    """Measurements of a batch that regrows during warm-up and measurement equal
    those of a batch created with room to spare (same seeds)."""
    prob = P.kanamori_2orb(beta=10.0, TaudiffMax=3.0, Nmeas=100, NCorr=40, Ntau=50, NLegMax=10)
    print("batched regrowth", flush=True)
    out = []
    for cap in (2, 64):
        gpu = CtqmcGpu.from_problem(prob, nchains=16, seed=7, capacity=cap)
        gpu.set_taudiffmax(3.0)
        gpu.warmup(2000)
        gpu.measure(100)
        out.append((gpu.accumulators(), gpu.chain_status()))
        gpu.close()
    (a, sa), (b, sb) = out
    assert sa["capacity"] > 2 and sb["capacity"] == 64
    assert np.array_equal(sa["kfmax"], sb["kfmax"])
    for k in ("gtau", "gleg", "histo", "outer_hist"):
        assert np.allclose(a[k], b[k], rtol=1e-9, atol=1e-12), k
    assert np.array_equal(a["scalars"][:, :12], b["scalars"][:, :12])


def test_fixed_capacity_failure_is_contained():
    """Chain 0 starts at the fixed capacity (4 pairs in one flavour): it stops
    with W2D_ERR_CAPACITY at its first move; the other chains complete and the
    pooled estimate leaves chain 0 out."""
    prob = P.kanamori_2orb(beta=2.0, TaudiffMax=-1.0, Nmeas=50, NCorr=20, Ntau=20, NLegMax=8)
    nch = 8
    gpu = CtqmcGpu.from_problem(prob, nchains=nch, seed=99, capacity=-4)
    taus = np.array([0.1, 0.3, 0.5, 0.7, 0.9, 1.1, 1.3, 1.5])
    cas = np.array([1, 2] * 4, np.int32)
    gpu.set_config(0, taus, np.zeros(8, np.int32), np.zeros(8, np.int32), cas, gpu.get_config(0)["outer_sst"])
    print("contained failure", flush=True)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        gpu.measure(50)
    st = gpu.chain_status()
    assert st["status"][0] == 3
    assert any("stopped on an error" in str(x.message) for x in w)
    others = st["status"][1:]
    assert np.all((others == 0) | (st["kfmax"][1:].max(1) >= 4))
    assert (others == 0).sum() >= 1
    agg = gpu.aggregate(ngroups=4)
    assert np.all(np.isfinite(agg["gtau"])) and np.isfinite(agg["mean_sign"])
    sums = gpu.group_sums(1)
    assert int(sums["nused"][0]) == int((st["status"] == 0).sum())
    gpu.close()


def test_abort_word_stops_at_a_measurement_boundary():
    prob = P.kanamori_2orb(beta=10.0, TaudiffMax=3.0, Nmeas=500, NCorr=20, Ntau=20, NLegMax=8)
    gpu = CtqmcGpu.from_problem(prob, nchains=4, seed=5)
    gpu.set_taudiffmax(3.0)
    gpu.warmup(500)
    print("abort", flush=True)
    gpu.set_abort(True)
    gpu.measure(500)
    assert gpu.aborted() == 1
    sc = gpu.accumulators()["scalars"]
    assert np.all(sc[:, 1] == 1)          # one measurement each, then stopped
    gpu.set_abort(False)
    gpu.measure(30)
    assert gpu.aborted() == 0
    assert np.all(gpu.accumulators()["scalars"][:, 1] == 30)
    gpu.close()
