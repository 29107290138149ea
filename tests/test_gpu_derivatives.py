"""GPU parameter derivatives against the oracle and exact results.

* d ln|det| / d theta and d ln|Tr| / d theta of oracle-visited configurations
  (GPU derivative kernel through the C-ABI vs the oracle, 1e-9 relative);
* score sums of a measurement run (same seeds, identical move sequences) vs
  the oracle's, 1e-9;
* score-function estimates d<n_up n_dn>/dtheta and dG(tau)/dtheta of batched
  chains vs exact diagonalisation of impurity + bath differentiated by central
  differences (3 sigma of the jackknife error).
"""

import numpy as np
import pytest

from oracle import OracleChain
from w2dynamics_amd import problems as P
from w2dynamics_amd.ctqmc import CtqmcGpu

pytestmark = pytest.mark.gpu

KAN0 = np.array([3.5, 0.6, 0.0, 1.0])
AND0 = np.array([2.0, -0.6, 1.0])


def _kan(th, **kw):
    p = dict(beta=8.0, TaudiffMax=-1.0, Ntau=20, NLegMax=5, nftau=801)
    p.update(kw)
    return P.kanamori_3orb_theta(th, **p)


def _and(th, **kw):
    p = dict(beta=5.0, Ntau=20, NLegMax=5, nftau=801)
    p.update(kw)
    return P.anderson_discrete_theta(th, **p)


MODELS = {"kanamori3": (_kan, KAN0), "anderson": (_and, AND0)}


@pytest.mark.parametrize("name", list(MODELS))
def test_dlogw_matches_oracle(name):
    mk, th0 = MODELS[name]
    prob, d = mk(th0)
    print("dlogw", name, flush=True)
    ora = OracleChain.from_problem(prob, 17)
    ora.set_derivatives(d, score=False)
    gpu = CtqmcGpu.from_problem(prob, nchains=1, seed=17)
    gpu.set_derivatives(d, score=False)
    for _ in range(4):
        ora.moves(1500, record=False)
        cfg = ora.config()
        gpu.set_config(0, cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
        a, b = gpu.dlogw(0), ora.dlogw()
        print("order", cfg["n"], a.ravel(), flush=True)
        assert np.allclose(a, b, rtol=1e-9, atol=1e-9), (a, b)
    gpu.close()


def _occ(dm, offsets, sizes, evecs):
    """[theta][state][state] eigenbasis blocks -> occupation basis (V rho V^T per block)."""
    out = np.array(dm, dtype=float, copy=True)
    for o, n, v in zip(offsets, sizes, evecs):
        out[..., o:o + n, o:o + n] = v @ out[..., o:o + n, o:o + n] @ v.T
    return out


@pytest.mark.parametrize("name", list(MODELS))
def test_score_sums_match_oracle(name):
    mk, th0 = MODELS[name]
    prob, d = mk(th0, NCorr=20, Nmeas=150)
    seed = 23
    print("score sums", name, flush=True)
    gpu = CtqmcGpu.from_problem(prob, nchains=2, seed=seed)
    gpu.set_derivatives(d)
    gpu.warmup(2000)
    gpu.measure(150)
    acc = gpu.accumulators()
    sums = gpu.score_sums(2)
    gm = gpu.model()
    for c in range(2):
        ora = OracleChain.from_problem(prob, seed + c)
        ora.set_derivatives(d)
        ora.warmup(2000)
        ora.measure(150, 20)
        ref = ora.score_sums()
        raw = ora.raw()
        assert np.allclose(acc["gtau"][c], raw["gtau"], rtol=1e-9, atol=1e-11)
        # density matrices in the occupation basis: each side's eigenvectors
        # are fixed only up to a gauge
        got = dict(sums)
        for k in ("dm", "dmx"):
            got[k] = np.array([_occ(sums[k][c], gm["offsets"], gm["sizes"], gm["evecs"])])
            ref[k] = _occ(ref[k], ora.model.offsets, ora.model.sizes, ora.model.evecs)
        for k in ("ssum", "gtau", "dm", "dmx"):
            a, b = np.asarray(got[k][0 if k in ("dm", "dmx") else c]), np.asarray(ref[k])
            assert a.shape == b.shape, k
            scale = max(1.0, np.abs(b).max())
            assert np.abs(a - b).max() <= 1e-9 * scale, (name, c, k, np.abs(a - b).max(), scale)
    gpu.close()


def _anderson_exact(th, beta, eps_k, v_k, ntau):
    from test_gpu_fullsize import _anderson_ed, _anderson_ed_gtau
    u, eps, v = th
    _, d = _anderson_ed(eps, u, np.asarray(eps_k), v * np.asarray(v_k), beta)
    g = _anderson_ed_gtau(eps, u, np.asarray(eps_k), v * np.asarray(v_k), beta, ntau)
    return d, g


def test_score_estimates_match_exact_diagonalisation():
    """Impurity + 3 bath sites, theta = (U, eps_d, V): d<n_up n_dn>/dtheta and
    dG(tau)/dtheta (bin averages) from 1024 chains within 3 sigma of central
    differences of exact diagonalisation."""
    beta, ntau = 5.0, 25
    eps_k, v_k = (-1.0, 0.2, 1.1), (0.5, 0.35, 0.45)
    prob, d = P.anderson_discrete_theta(AND0, beta=beta, eps_k=eps_k, v_k=v_k, nftau=4001, Ntau=ntau,
                                        NCorr=20, Nmeas=2000)
    gpu = CtqmcGpu.from_problem(prob, nchains=1024, seed=4100)
    gpu.set_derivatives(d)
    gpu.warmup(5000)
    gpu.measure(2000)
    est = gpu.score_estimate(ngroups=64)
    gpu.close()
    h = 1e-4
    for t, name in enumerate(P.ANDERSON_THETA):
        e = np.zeros(3)
        e[t] = h
        dp, gp = _anderson_exact(AND0 + e, beta, eps_k, v_k, ntau)
        dm, gm = _anderson_exact(AND0 - e, beta, eps_k, v_k, ntau)
        dd, dg = (dp - dm) / (2 * h), (gp - gm) / (2 * h)
        val, err = est["ddouble_occ"][t][0], est["ddouble_occ_err"][t][0]
        print(name, "d docc", val, "+-", err, "exact", dd, flush=True)
        assert abs(val - dd) <= 3 * err + 5e-4, (name, val, err, dd)
        # the reference's G(tau) is -<T c(tau) c+> with its sign reversed
        g, ge = est["dgtau"][t][0, 0], est["dgtau_err"][t][0, 0]
        dev = np.abs(g + dg) / (ge + 5e-4)
        print(name, "dG largest deviation in sigma", dev.max(), flush=True)
        assert np.mean(dev <= 3) >= 0.95 and dev.max() <= 4.5, (name, dev.max())


def test_autograd_gradients_match_crn_finite_differences():
    """3-orbital Kanamori, theta = (U, J, eps_d, V), 1024 chains: gradients that
    the torch.autograd binding back-propagates (score-function Jacobian) for
    the double occupancies (w.r.t. U) and G(tau) (w.r.t. J) against central
    differences with common random numbers (same seeds, same warm start):
    within 3 sigma at delta = 0.05 and within 2 sigma at delta = 1e-3, the
    latter dominated by its own noise."""
    import torch
    from w2dynamics_amd.autograd import ScoreRun, observables
    from w2dynamics_amd.derivatives import finite_difference_crn

    def builder(th):
        return P.kanamori_3orb_theta(th, beta=10.0, Ntau=20, NLegMax=10, NCorr=20, Nmeas=1000, nftau=1001)

    nchains, seed = 1024, 777
    run = ScoreRun(builder, nchains=nchains, seed=seed, nwarm=4000, nmeas=1000)
    theta = torch.tensor(KAN0, dtype=torch.float64, requires_grad=True)
    g, d = observables(theta, run)
    est = run.last
    # gradient of the summed double occupancy: the autograd path against the estimate
    d.sum().backward()
    assert np.allclose(theta.grad.numpy(), est["ddouble_occ"].sum(1), rtol=1e-12, atol=1e-12)
    for t, key, delta, nsig in ((0, "double_occ", 0.05, 3.0), (1, "gtau", 0.05, 3.0), (0, "double_occ", 1e-3, 2.0)):
        fd = finite_difference_crn(builder, KAN0, t, delta, run.configs, nchains, seed + 1, 1000)
        sc, se = est["d" + key][t], est["d" + key + "_err"][t]
        dev = np.abs(sc - fd[key]) / np.sqrt(se ** 2 + fd[key + "_err"] ** 2 + 1e-30)
        print(P.KANAMORI_THETA[t], key, "delta", delta, "score", np.ravel(sc)[:3], "fd", np.ravel(fd[key])[:3],
              "largest deviation in sigma", dev.max(), flush=True)
        assert np.mean(dev <= nsig) >= 0.9 and dev.max() <= nsig + 1.5, (t, key, delta, dev.max())
