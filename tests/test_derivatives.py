"""Parameter derivatives on the CPU: the input builders, the oracle's
d ln|w| / d theta and its score sums against finite differences at fixed
configurations, and the score-function estimator algebra against an exactly
weighted ensemble.  (No counterpart in the reference: these pin the oracle
the GPU derivative path is compared with.)"""

import numpy as np
import pytest

from oracle import OracleChain
from w2dynamics_amd import problems as P
from w2dynamics_amd.derivatives import score_estimate

KAN0 = np.array([3.5, 0.6, 0.0, 1.0])
AND0 = np.array([2.0, -0.6, 1.0])


def _kan(th):
    return P.kanamori_3orb_theta(th, beta=8.0, TaudiffMax=-1.0, Ntau=20, NLegMax=5, nftau=801)


def _and(th):
    return P.anderson_discrete_theta(th, beta=5.0, Ntau=20, NLegMax=5, nftau=801)


MODELS = {"kanamori3": (_kan, KAN0), "anderson": (_and, AND0)}


@pytest.mark.parametrize("name", list(MODELS))
def test_input_derivatives_match_finite_differences(name):
    mk, th0 = MODELS[name]
    _, d = mk(th0)
    for t in range(len(th0)):
        e = np.zeros(len(th0))
        e[t] = 1e-3
        pp, _ = mk(th0 + e)
        pm, _ = mk(th0 - e)
        # U, muimp linear and F ~ V^2: central differences are exact up to rounding
        for a, b, da in ((pp.umatrix, pm.umatrix, d.du[t]), (pp.muimp, pm.muimp, d.dmuimp[t]),
                         (pp.ftau, pm.ftau, d.dftau[t])):
            assert np.abs((a - b) / 2e-3 - da).max() <= 1e-9 * max(1.0, np.abs(da).max())


def _configs(mk, th0, seed, n=3, moves=1500):
    prob, _ = mk(th0)
    o = OracleChain.from_problem(prob, seed)
    out = []
    for _ in range(n):
        o.moves(moves, record=False)
        out.append(o.config())
    return out


def _logw(mk, th, cfg):
    prob, _ = mk(th)
    o = OracleChain.from_problem(prob, 1)
    w = o.set_config(cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
    # trace_log carries beta * Egs (weights relative to the ground state)
    return w["det_log"], w["trace_log"] - prob.beta * o.model.egs


@pytest.mark.parametrize("name", list(MODELS))
def test_oracle_dlogw_matches_finite_differences(name):
    """d ln|det| and d ln|Tr| of visited configurations against central
    differences of the oracle's own weights (models rebuilt at theta +- h)."""
    mk, th0 = MODELS[name]
    prob, d = mk(th0)
    h = 1e-4
    for cfg in _configs(mk, th0, 5):
        o = OracleChain.from_problem(prob, 1)
        o.set_derivatives(d, score=False)
        o.set_config(cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
        dl = o.dlogw()
        for t in range(len(th0)):
            e = np.zeros(len(th0))
            e[t] = h
            wp, wm = _logw(mk, th0 + e, cfg), _logw(mk, th0 - e, cfg)
            fd = [(wp[i] - wm[i]) / (2 * h) for i in range(2)]
            for i in range(2):
                assert abs(dl[t, i] - fd[i]) <= 1e-6 * max(1.0, abs(fd[i])), (name, cfg["n"], t, i, dl[t, i], fd[i])
        # V enters only the determinant: d ln|det| / dV = 2 k / V
        if name == "anderson":
            assert abs(dl[2, 0] - cfg["n"] / th0[2]) <= 1e-9 * cfg["n"]


def _occ_dm(o):
    r = o.raw()
    m = o.model
    dm = r["dm"].copy()
    for s in range(m.nsst):
        a, n = m.offsets[s], m.sizes[s]
        dm[a:a + n, a:a + n] = m.evecs[s] @ dm[a:a + n, a:a + n] @ m.evecs[s].T
    return dm, r


@pytest.mark.parametrize("name", list(MODELS))
def test_oracle_score_sums_of_one_measurement(name):
    """One measurement on a fixed configuration: Sum sign s = sign x (d ln|det| +
    d ln|Tr|), the G(tau) and density-matrix products, and the explicit
    derivative of the density-matrix sample against central differences of the
    sample in the occupation basis."""
    mk, th0 = MODELS[name]
    prob, d = mk(th0)
    h = 1e-5
    for cfg in _configs(mk, th0, 11, n=2, moves=2000):
        o = OracleChain.from_problem(prob, 1)
        o.set_derivatives(d, score=True)
        o.set_config(cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
        o.measure(1, 0)
        sc, s = o.score_sums(), o.dlogw().sum(1)
        _, raw = _occ_dm(o)
        assert np.allclose(sc["ssum"], raw["scalars"][0] * s, rtol=1e-12, atol=1e-12)
        assert np.allclose(sc["gtau"], s[:, None, None] * raw["gtau"][None], rtol=1e-12, atol=1e-14)
        assert np.allclose(sc["dm"], s[:, None, None] * raw["dm"][None], rtol=1e-12, atol=1e-14)
        m = o.model
        for t in range(len(th0)):
            e = np.zeros(len(th0))
            e[t] = h
            vals = []
            for th in (th0 + e, th0 - e):
                pr, _ = mk(th)
                oo = OracleChain.from_problem(pr, 1)
                oo.set_config(cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
                oo.measure(1, 0)
                vals.append(_occ_dm(oo)[0])
            fd = (vals[0] - vals[1]) / (2 * h)
            dmx = sc["dmx"][t].copy()
            for k in range(m.nsst):
                a, n = m.offsets[k], m.sizes[k]
                dmx[a:a + n, a:a + n] = m.evecs[k] @ dmx[a:a + n, a:a + n] @ m.evecs[k].T
            assert np.abs(dmx - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max()), (name, t)


def test_score_estimator_on_an_exactly_weighted_ensemble():
    """score_estimate on sums that are exact ensemble averages: K
    'configurations' with signs, weights |w_k| exp(theta a_k), G-samples whose
    bath dependence scales like exp(-lambda theta) and density-matrix samples
    with an explicit derivative.  Its derivatives equal central differences of
    the normalised observables of the reweighted ensemble."""
    rng = np.random.default_rng(3)
    nb, ntau, beta, nst = 1, 6, 4.0, 4
    K = 40
    sgn = np.where(rng.random(K) < 0.85, 1.0, -1.0)
    a = rng.normal(size=(2, K))                 # d ln|w_k| / d theta_t
    w0 = rng.random(K) + 0.2
    g0 = rng.normal(size=(K, 2, ntau))
    lam = np.array([0.0, 0.7])                  # G explicit: g_k(theta_1) = g_k exp(-0.7 theta_1)
    dmk = rng.normal(size=(K, nst, nst))
    dmk = dmk + dmk.transpose(0, 2, 1)
    ddm = rng.normal(size=(2, K, nst, nst))
    ddm = ddm + ddm.transpose(0, 1, 3, 2)
    model = dict(offsets=np.array([0]), sizes=np.array([nst]), evecs=[np.linalg.qr(rng.normal(size=(nst, nst)))[0]],
                 states=np.arange(nst))
    # the G estimator (elements of Delta^-1) scales like exp(-lambda theta)
    # when F does like exp(+lambda theta): dF/dtheta_1 = 0.7 F
    ftau = np.ones((2, 5))
    dftau = np.stack([np.zeros((2, 5)), lam[1] * ftau])

    def sums_at(th):
        p = w0 * np.exp(a.T @ th)
        g = g0 * np.exp(-lam[1] * th[1])
        dm = dmk + np.einsum("t,tkij->kij", th, ddm)
        return dict(gtau=[np.einsum("k,k,kft->ft", p, sgn, g)], dm=[np.einsum("k,k,kij->ij", p, sgn, dm)],
                    scalars=[np.array([np.sum(p * sgn), np.sum(p)] + [1.0] * 12)]), p, g, dm

    def normalised(th):
        s, _, _, _ = sums_at(th)
        S = s["scalars"][0][0]
        g = s["gtau"][0] / S * ntau / beta ** 2
        dm = -s["dm"][0] / S
        v = model["evecs"][0]
        return g, v @ dm @ v.T

    th0 = np.zeros(2)
    s, p, g, dm = sums_at(th0)
    score = dict(ssum=[np.einsum("k,k,tk->t", p, sgn, a)], gtau=[np.einsum("k,k,uk,kft->uft", p, sgn, a, g)],
                 dm=[np.einsum("k,k,tk,kij->tij", p, sgn, a, dm)], dmx=[np.einsum("k,k,tkij->tij", p, sgn, ddm)])
    # one group twice (jackknife needs two): half of every sum in each
    half = lambda d: {k: [np.asarray(v[0]) / 2] * 2 for k, v in d.items()}  # noqa: E731
    est = score_estimate(half(s), half(score), nb, beta, ntau, model, ftau, dftau)
    for t in range(2):
        e = np.zeros(2)
        e[t] = 1e-6
        gp, dp = normalised(th0 + e)
        gm, dmm = normalised(th0 - e)
        fdg = (gp - gm) / 2e-6
        fdd = (dp - dmm) / 2e-6
        assert np.allclose(est["dgtau"][t].swapaxes(0, 1).reshape(2, ntau), fdg, rtol=1e-6, atol=1e-8)
        states = model["states"]
        nocc = np.array([[(states >> (b + nb * s_)) & 1 for s_ in range(2)] for b in range(nb)])
        docc = np.einsum("k,ijk,lmk->ijlm", np.diagonal(fdd), nocc, nocc)
        assert np.allclose(est["docc"][t], docc, rtol=1e-6, atol=1e-8)


def test_autograd_function_applies_the_score_jacobian():
    """CtqmcObservables: forward returns the run's estimates, backward the
    vector-Jacobian product with dG/dtheta and d(double occupancy)/dtheta;
    NaN Jacobian entries only matter where the incoming gradient is non-zero."""
    import torch
    from w2dynamics_amd.autograd import observables
    rng = np.random.default_rng(5)
    nb, ntau, nt = 2, 7, 3
    est = dict(gtau=rng.normal(size=(nb, 2, ntau)), double_occ=rng.normal(size=nb),
               dgtau=rng.normal(size=(nt, nb, 2, ntau)), ddouble_occ=rng.normal(size=(nt, nb)))
    est["dgtau"][2, 1] = np.nan                  # not provided for theta_2, orbital 1

    seen = []

    def run(theta):
        seen.append(np.array(theta))
        return est

    theta = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, requires_grad=True)
    g, d = observables(theta, run)
    assert np.allclose(seen[0], [1.0, 2.0, 3.0])
    assert np.allclose(g.detach().numpy(), est["gtau"]) and np.allclose(d.detach().numpy(), est["double_occ"])
    w = rng.normal(size=(nb, 2, ntau))
    w[1] = 0.0
    loss = (g * torch.as_tensor(w)).sum() + 3.0 * d.sum()
    loss.backward()
    dg = np.nan_to_num(est["dgtau"])
    want = np.einsum("tbsk,bsk->t", dg, w) + 3.0 * est["ddouble_occ"].sum(1)
    assert np.allclose(theta.grad.numpy(), want, rtol=1e-12, atol=1e-12)
