"""GPU: G(i w_n) measured from M^-1 (MeasGiw, measure_giw_nfft, CTQMC.F90:1015-1059)
against the oracle's direct sum on identical chains, and against the Matsubara
transform of the same run's Legendre estimator (exact by the plane-wave
expansion) on a batch of chains.
"""

import numpy as np
import pytest

from oracle import OracleChain
from test_oracle import legendre_to_matsubara
from w2dynamics_amd import problems as P
from w2dynamics_amd.ctqmc import CtqmcError, CtqmcGpu

W2D_ERR_STATE = 6

pytestmark = pytest.mark.gpu

# Niw = 1100: two frequency passes of a 256-thread workgroup, the second partial
CASES = {
    "c1": (P.anderson_1orb(beta=10.0, MeasGiw=1, Niw=1100, Nmeas=100, NCorr=40, Ntau=100, NLegMax=20), None),
    "c2_win": (P.kanamori_2orb(beta=10.0, MeasGiw=1, Niw=1100, Nmeas=100, NCorr=40, Ntau=100, NLegMax=20,
                               TaudiffMax=3.0), 3.0),
}


def _run_pair(prob, tdm, seed, nchains, nwarm, nmeas, ncorr):
    gpu = CtqmcGpu.from_problem(prob, nchains=nchains, seed=seed)
    if tdm:
        gpu.set_taudiffmax(tdm)
    gpu.warmup(nwarm)
    gpu.measure(nmeas)
    res = gpu.results()
    orc = []
    for c in range(nchains):
        ora = OracleChain.from_problem(prob, seed + c)
        if tdm:
            ora.set_taudiffmax(tdm)
        ora.warmup(nwarm)
        ora.measure(nmeas, ncorr)
        orc.append(ora)
    return gpu, res, orc


@pytest.mark.parametrize("case", list(CASES))
def test_giw_matches_oracle(case):
    prob, tdm = CASES[case]
    print("giw vs oracle", case, flush=True)
    gpu, res, orc = _run_pair(prob, tdm, 2024, 2, 3000, 100, 40)
    for c, ora in enumerate(orc):
        r = ora.results()
        # the chain itself is unchanged by the launch-per-measurement schedule
        assert np.allclose(res["gtau"][c], r["gtau"], rtol=1e-9, atol=1e-11)
        a, b = res["giw"][c], r["giw"]
        assert a.shape == b.shape == (prob.nbands, 2, 2 * 1100)
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), (c, np.abs(a - b).max())
    gpu.close()


def test_giw_matches_oracle_at_orders_above_one_tile():
    """Flavours with more than 64 pairs (several annihilator and creator tiles)."""
    prob = P.anderson_1orb(beta=500.0, u=0.5, nftau=4001, MeasGiw=1, Niw=300, Nmeas=20, NCorr=200,
                           Ntau=200, NLegMax=10)
    print("giw vs oracle, large orders", flush=True)
    gpu, res, orc = _run_pair(prob, None, 5, 1, 60000, 20, 200)
    kmax = gpu.chain_status()["kfmax"][0]
    assert max(kmax) > 64, kmax
    a, b = res["giw"][0], orc[0].results()["giw"]
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), np.abs(a - b).max()
    gpu.close()


def test_giw_is_the_matsubara_transform_of_gleg_on_a_batch():
    """2-orbital Kanamori at beta = 50 on 256 chains: every chain's G(i w_n)
    equals sum_l T_nl G_l of the same measurements (NLegMax = 160 leaves a
    truncation below 1e-20 for |w_n| beta / 2 <= 99), and so do the pooled
    estimates."""
    niw, nleg = 32, 160
    prob = P.kanamori_2orb(beta=50.0, MeasGiw=1, Niw=niw, NLegMax=nleg, Nmeas=50, NCorr=100, Ntau=200)
    print("giw vs Legendre transform, 256 chains", flush=True)
    gpu = CtqmcGpu.from_problem(prob, nchains=256, seed=77)
    gpu.warmup(20000)
    gpu.measure(50)
    res = gpu.results()
    T = legendre_to_matsubara(nleg, niw)
    want = np.einsum("nl,cbsl->cbsn", T, res["gleg"])
    scale = np.abs(res["giw"]).max()
    assert np.abs(res["giw"] - want).max() <= 1e-10 * scale
    agg = gpu.aggregate(16)
    pooled = np.einsum("nl,bsl->bsn", T, agg["gleg"])
    assert np.abs(agg["giw"] - pooled).max() <= 1e-10 * np.abs(pooled).max()
    assert np.isfinite(agg["giw_err"]).all() and (agg["giw_err"] > 0).all()
    # G(i w) of a real estimator: negative frequencies are the conjugates
    assert np.array_equal(res["giw"][..., :niw], np.conj(res["giw"][..., niw:][..., ::-1]))
    gpu.close()


def test_giw_off_by_default():
    prob = P.kanamori_2orb(beta=10.0, Nmeas=10, NCorr=20)
    gpu = CtqmcGpu.from_problem(prob, nchains=2, seed=1)
    assert gpu.niw == 0
    gpu.warmup(100)
    gpu.measure(10)
    assert "giw" not in gpu.results()
    with pytest.raises(CtqmcError) as e:
        gpu.giw_sums(0)
    assert e.value.code == W2D_ERR_STATE
    gpu.close()
