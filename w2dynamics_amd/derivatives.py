"""Score-function derivatives of CT-HYB observables with respect to model
parameters theta (no counterpart in the reference).

With configuration weights w(C; theta) (the reference's det(Delta) x local
trace, ``MatrixUpdate.F90:41-99``, ``Trace.F90:5734-6102``) and signed
Monte Carlo averages <X> = Sum sign X / Sum sign,

    d<O>/dtheta = <O s> - <O><s> + <dO/dtheta>,   s = d ln|w| / dtheta,

where the last term is the explicit dependence of the estimator itself.  The
engine accumulates, per chain, Sum sign s, Sum s x (G(tau) sample), Sum s x
(density-matrix sample) and Sum d(density-matrix sample)/dtheta at fixed
configuration (``CtqmcGpu.score_sums``).  The G(tau) estimator depends on theta
only through Delta: where dF/dtheta = lambda_f F_f for every flavour (no bath
dependence: lambda = 0; a bath coupling V with F ~ V^2: lambda = 2/V), its
explicit term is -lambda_f <G_f>; for other dF/dtheta it is not formed and
dG/dtheta is NaN.
"""

from __future__ import annotations

import numpy as np


def _proportionality(dftau: np.ndarray, ftau: np.ndarray, rtol: float = 1e-12) -> np.ndarray:
    """lambda[t, f] with dF_t,f = lambda F_f (0 if dF vanishes, NaN if not proportional)."""
    nt, nf = dftau.shape[:2]
    lam = np.full((nt, nf), np.nan)
    for t in range(nt):
        for f in range(nf):
            d, F = dftau[t, f], ftau[f]
            scale = np.abs(d).max()
            if scale == 0.0:
                lam[t, f] = 0.0
                continue
            x = float(d @ F / (F @ F))
            if np.abs(d - x * F).max() <= rtol * max(scale, 1e-300):
                lam[t, f] = x
    return lam


def _estimate(tot: dict, sc: dict, nb: int, beta: float, ntau: int, model: dict, lam: np.ndarray) -> dict:
    S = tot["scalars"][0]
    nt = sc["ssum"].shape[0]
    ss = sc["ssum"] / S                                   # <s>
    nf = 2 * nb

    def fl2(a):   # [..., f, x] -> [..., b, s, x]
        return a.reshape(a.shape[:-2] + (2, nb, a.shape[-1])).swapaxes(-3, -2)

    g = tot["gtau"] / S * ntau / beta ** 2                # [f][tau]
    dg = (sc["gtau"] / S * ntau / beta ** 2 - g[None] * ss[:, None, None]
          - lam[:, :, None] * g[None])                    # [t][f][tau]
    dm = -tot["dm"] / S
    ddm = -(sc["dm"] / S - tot["dm"][None] / S * ss[:, None, None] + sc["dmx"] / S)
    dm_occ = dm.copy()
    ddm_occ = ddm.copy()
    for k, (o, n) in enumerate(zip(model["offsets"], model["sizes"])):
        v = model["evecs"][k]
        dm_occ[o:o + n, o:o + n] = v @ dm[o:o + n, o:o + n] @ v.T
        ddm_occ[:, o:o + n, o:o + n] = v @ ddm[:, o:o + n, o:o + n] @ v.T
    states = np.asarray(model["states"])
    nocc = np.array([[(states >> (b + nb * s)) & 1 for s in range(2)] for b in range(nb)])
    occ = np.einsum("k,ijk,lmk->ijlm", np.diagonal(dm_occ), nocc, nocc)
    docc_t = np.einsum("tk,ijk,lmk->tijlm", np.diagonal(ddm_occ, axis1=1, axis2=2), nocc, nocc)
    b = np.arange(nb)
    return dict(gtau=fl2(g), dgtau=fl2(dg), occ=occ, docc=docc_t,
                double_occ=occ[b, 0, b, 1], ddouble_occ=docc_t[:, b, 0, b, 1], dlogw_mean=ss,
                mean_sign=S / tot["scalars"][1])


def score_estimate(sums: dict, score: dict, nb: int, beta: float, ntau: int, model: dict,
                   ftau: np.ndarray, dftau: np.ndarray) -> dict:
    """Pooled observables and score-function derivatives over the leading
    (group) axis of ``sums`` (``CtqmcGpu.group_sums``) and ``score``
    (``CtqmcGpu.score_sums``), with jackknife errors over that axis
    (``<key>_err``).  ftau [f][tau] and dftau [theta][f][tau] are the
    hybridisation tables per flavour f = b + N s.

    Keys: gtau [b][s][tau], dgtau [theta][b][s][tau] (NaN where dF/dtheta is
    not proportional to F), occ [b][s][b][s] and docc [theta][...],
    double_occ [b] = <n_b,up n_b,down> and ddouble_occ [theta][b],
    dlogw_mean [theta] = <d ln|w| / dtheta>."""
    lam = _proportionality(np.asarray(dftau, float), np.asarray(ftau, float))
    keys = ("gtau", "dm", "scalars")
    skeys = ("ssum", "gtau", "dm", "dmx")
    tot = {k: np.asarray(sums[k], float).sum(0) for k in keys}
    stot = {k: np.asarray(score[k], float).sum(0) for k in skeys}
    out = _estimate(tot, stot, nb, beta, ntau, model, lam)
    n = len(sums["scalars"])
    if n < 2:
        for k in list(out):
            out[k + "_err"] = np.full_like(np.asarray(out[k], float), np.nan)
        return out
    jk = {k: [] for k in out}
    for i in range(n):
        t = {k: tot[k] - np.asarray(sums[k], float)[i] for k in keys}
        st = {k: stot[k] - np.asarray(score[k], float)[i] for k in skeys}
        r = _estimate(t, st, nb, beta, ntau, model, lam)
        for k in out:
            jk[k].append(np.asarray(r[k], float))
    for k in list(out):
        a = np.array(jk[k])
        out[k + "_err"] = np.sqrt((n - 1) / n * ((a - a.mean(0)) ** 2).sum(0))
    return out


def _pooled(sums: dict, nb: int, beta: float, ntau: int, model: dict) -> dict:
    """Sign-ratio estimates of G(tau) and the double occupancy from group sums,
    each group left out in turn (row 0: all groups; row 1 + i: without group i)."""
    keys = ("gtau", "dm", "scalars")
    tot = {k: np.asarray(sums[k], float).sum(0) for k in keys}
    n = len(sums["scalars"])
    rows = [tot] + [{k: tot[k] - np.asarray(sums[k], float)[i] for k in keys} for i in range(n)]
    out = []
    for r in rows:
        S = r["scalars"][0]
        g = r["gtau"] / S * ntau / beta ** 2
        dm = -r["dm"] / S
        diag = np.zeros(dm.shape[0])
        for k, (o, m_) in enumerate(zip(model["offsets"], model["sizes"])):
            v = model["evecs"][k]
            diag[o:o + m_] = np.einsum("ij,jk,ik->i", v, dm[o:o + m_, o:o + m_], v)
        states = np.asarray(model["states"])
        d = np.array([np.sum(diag * ((states >> b) & 1) * ((states >> (b + nb)) & 1)) for b in range(nb)])
        out.append(dict(gtau=g.reshape(2, nb, ntau).swapaxes(0, 1), double_occ=d))
    return out


def finite_difference_crn(builder, theta, t: int, delta: float, configs, nchains: int, seed: int,
                          nmeas: int, taudiffmax: float | None = None, capacity: int = 256,
                          ngroups: int = 64, device: int = 0, engine: int = 0) -> dict:
    """Central difference (O(theta + delta e_t) - O(theta - delta e_t)) / (2 delta)
    of G(tau) and the double occupancy with common random numbers: both runs
    start every chain from the same configuration (``configs``, as
    ``CtqmcGpu.get_config`` returns them, e.g. thermalised at theta) with the
    same seeds, so that their Markov chains coincide until an acceptance
    decision differs.  Errors: jackknife over groups of chains applied to the
    difference (which keeps the correlation of the paired runs)."""
    from .ctqmc import CtqmcGpu
    res = []
    for sgn in (1.0, -1.0):
        th = np.array(theta, float)
        th[t] += sgn * delta
        prob, _ = builder(th)
        gpu = CtqmcGpu.from_problem(prob, nchains=nchains, seed=seed, capacity=capacity, device=device)
        if engine > 0:
            gpu.set_engine(engine)
        if taudiffmax is not None:
            gpu.set_taudiffmax(taudiffmax)
        for c, cfg in enumerate(configs[:nchains]):
            gpu.set_config(c, cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"], cfg["outer_sst"])
        gpu.measure(nmeas)
        ok = int((gpu.chain_status()["status"] == 0).sum())
        if ok != nchains:
            raise RuntimeError("finite_difference_crn: %d of %d chains stopped" % (nchains - ok, nchains))
        res.append(_pooled(gpu.group_sums(int(ngroups)), prob.nbands, prob.beta, gpu.ntau, gpu.model()))
        gpu.close()
    out = {}
    n = len(res[0]) - 1
    for k in ("gtau", "double_occ"):
        vals = np.array([(p[k] - m[k]) / (2 * delta) for p, m in zip(res[0], res[1])])
        jk = vals[1:]
        out[k] = vals[0]
        out[k + "_err"] = np.sqrt((n - 1) / n * ((jk - jk.mean(0)) ** 2).sum(0))
    return out
