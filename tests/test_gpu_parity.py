"""GPU parity: the HIP engine against the CPU oracle (and the reference's golden
vectors) on the same seeded inputs.

Decisions (move kind, accept/reject, order, outer superstate, window position)
must be identical move by move; log-weights agree to 1e-10 absolute (i.e. the
weights to ~1e-10 relative) as the sampling path is FP64 throughout.
"""

import numpy as np
import pytest

from conftest import load_golden, qmc_of
from oracle import OracleChain
from w2dynamics_amd import problems as P
from w2dynamics_amd.ctqmc import CtqmcGpu

pytestmark = pytest.mark.gpu

TOL_LOG = 1e-10
FIELDS = ("kind", "outcome", "noper", "outer_sst", "window")


def _compare(rec, ref, label):
    for f in FIELDS:
        bad = np.nonzero(rec[f] != ref[f])[0]
        assert len(bad) == 0, "%s: %s differs first at move %d (gpu %s, oracle %s)" % (
            label, f, bad[0], rec[f][bad[0]], ref[f][bad[0]])
    for f in ("trace_log", "det_log"):
        d = np.abs(rec[f] - ref[f])
        assert d.max() <= TOL_LOG, "%s: %s differs by %.3e" % (label, f, d.max())
    for f in ("trace_sign", "det_sign", "cfgsign"):
        assert np.array_equal(rec[f], ref[f]), "%s: %s differs" % (label, f)


CASES = {
    "c1_w1": (lambda: P.anderson_1orb(), 43890, 6000),
    "c1_win": (lambda: P.anderson_1orb(TaudiffMax=2.0), 4711, 6000),
    "c2_win": (lambda: P.kanamori_2orb(beta=20.0, TaudiffMax=4.0), 12345, 6000),
    "c3_win": (lambda: P.kanamori_3orb(beta=20.0, TaudiffMax=5.0), 777, 3000),
    "c4_w1": (lambda: P.coulomb_5orb(beta=2.0), 99, 600),
    # 5-orbital blocks (FP64 MFMA walks, HBM model tables) with a sliding window
    "c4_win": (lambda: P.coulomb_5orb(beta=4.0, TaudiffMax=1.0), 2024, 800),
}


@pytest.mark.parametrize("case", list(CASES))
def test_move_log_matches_oracle(case):
    mk, seed, n = CASES[case]
    prob = mk()
    tdm = float(prob.qmc()["TaudiffMax"])
    print("case", case, "seed", seed, "moves", n, flush=True)
    gpu = CtqmcGpu.from_problem(prob, nchains=2, seed=seed)
    gpu.set_taudiffmax(tdm)
    rec = gpu.moves_log(0, n)
    ora = OracleChain.from_problem(prob, seed)
    ora.set_taudiffmax(tdm)
    ref = ora.moves(n)
    _compare(rec, ref, case)
    # the final inverse hybridisation blocks agree as well
    for f in range(2 * prob.nbands):
        a, b = gpu.get_minv(0, f), ora.minv(f)
        assert a.shape == b.shape
        if a.size:
            assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(b).max())
    gpu.close()


@pytest.mark.parametrize("name", ["c1_log_w1", "c2_log_win", "c3_log_win", "c3_log_all", "c3d_log_all", "c4_log_all",
                                  "c2_log_sym", "c3_log_sym"])
def test_move_log_matches_reference_golden(name):
    z = load_golden(name)
    q = qmc_of(z)
    prob = P.Problem(int(z["nbands"]), float(z["beta"]), z["umatrix"], z["ftau"], z["muimp"],
                     params={k: v for k, v in q.items()})
    print("golden", name, flush=True)
    gpu = CtqmcGpu(str(z["param_string"]), z["umatrix"], z["ftau"], z["muimp"], nchains=1,
                   seed=int(z["seed"]))
    gpu.set_taudiffmax(float(q["TaudiffMax"]))
    n = int(z["nmoves"])          # the whole reference log
    rec = gpu.moves_log(0, n)
    ref = {f: z["move_" + f][:n] for f in FIELDS + ("trace_log", "det_log", "trace_sign",
                                                    "det_sign", "cfgsign")}
    _compare(rec, ref, name)
    gpu.close()


def test_batched_chains_are_independent_streams():
    prob = P.kanamori_2orb(beta=10.0, TaudiffMax=3.0)
    seed, n = 2024, 1500
    print("batched chains", flush=True)
    gpu = CtqmcGpu.from_problem(prob, nchains=3, seed=seed)
    gpu.set_taudiffmax(3.0)
    for c in range(3):
        rec = gpu.moves_log(c, n)
        ora = OracleChain.from_problem(prob, seed + c)
        ora.set_taudiffmax(3.0)
        _compare(rec, ora.moves(n), "chain %d" % c)
    gpu.close()


def test_measurement_run_matches_oracle():
    prob = P.kanamori_2orb(beta=10.0, TaudiffMax=3.0, Nmeas=200, NCorr=40, Ntau=100, NLegMax=20)
    seed = 2024
    print("measurement run", flush=True)
    gpu = CtqmcGpu.from_problem(prob, nchains=2, seed=seed)
    gpu.set_taudiffmax(3.0)
    gpu.warmup(3000)
    gpu.measure(200)
    from w2dynamics_amd.ctqmc import check_accumulators
    check_accumulators(gpu.accumulators())
    res = gpu.results()
    for c in range(2):
        ora = OracleChain.from_problem(prob, seed + c)
        ora.set_taudiffmax(3.0)
        ora.warmup(3000)
        ora.measure(200, 40)
        r = ora.results()
        for k in ("mean_sign", "gtau", "gleg", "occ", "densitymatrix", "histo", "acc",
                  "outer_sst_histo"):
            a, b = np.asarray(res[k][c]), np.asarray(r[k])
            assert a.shape == b.shape, k
            assert np.allclose(a, b, rtol=1e-9, atol=1e-11, equal_nan=True), (c, k, np.abs(a - b).max())
    gpu.close()


def test_fixed_configurations_weights():
    """set_config on configurations visited by the oracle: trace, determinant,
    sign and inverse blocks."""
    prob = P.kanamori_3orb(beta=20.0, TaudiffMax=5.0)
    ora = OracleChain.from_problem(prob, 31)
    ora.set_taudiffmax(5.0)
    gpu = CtqmcGpu.from_problem(prob, nchains=1, seed=31)
    for step in range(4):
        ora.moves(700, record=False)
        cfg = ora.config()
        w = gpu.set_config(0, cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
        ref = ora.set_config(cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
        print("config", step, "order", cfg["n"], flush=True)
        assert abs(w["trace_log"] - ref["trace_log"]) <= TOL_LOG
        assert abs(w["det_log"] - ref["det_log"]) <= TOL_LOG
        for k in ("trace_sign", "det_sign", "cfgsign"):
            assert w[k] == ref[k], k
    gpu.close()
