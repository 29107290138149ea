"""Statistics of bench.py's G(tau) efficiency figures (host only)."""

import importlib.util
import os

import numpy as np

from conftest import ROOT


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _acc(rng, nchains, signs):
    """Raw per-chain sums as the engine returns them: 40 measurements per chain,
    G samples of mean -0.3 with unit noise, signs as given."""
    nmeas, nf, ntau = 40, 2, 8
    s = signs(rng, (nchains, nmeas))
    g = (-0.3 + rng.standard_normal((nchains, nmeas, nf, ntau))) * s[:, :, None, None]
    sc = np.zeros((nchains, 14))
    sc[:, 0] = s.sum(1)
    sc[:, 1] = nmeas
    return {"gtau": g.sum(1), "scalars": sc}


def _groups(acc, ngroups):
    """Sums over contiguous groups of chains, as the GPU's group sums."""
    idx = np.array_split(np.arange(acc["gtau"].shape[0]), ngroups)
    return (np.stack([acc["gtau"][i].sum(0) for i in idx]),
            np.array([acc["scalars"][i, 0].sum() for i in idx]))


def test_group_ratio_estimator_matches_per_chain_without_sign_problem():
    b = _bench()
    rng = np.random.default_rng(7)
    acc = _acc(rng, 1024, lambda r, sh: np.ones(sh))
    ntau, beta = 8, 2.0
    per_chain = acc["gtau"] / acc["scalars"][:, 0][:, None, None] * ntau / beta ** 2
    ref = b.gtau_inv_err2({"gtau": per_chain})
    g, s = _groups(acc, 64)
    blk = b.ratio_inv_err2(g, s, ntau, beta)
    # same expectation; 64 group samples give a ~18% standard deviation of 1/err^2
    assert 0.6 < blk / ref < 1.6


def test_jackknife_equals_block_standard_error_without_sign_problem():
    """For a linear estimator (all signs +1, equal block counts) the jackknife
    error is the standard error of the block means exactly."""
    b = _bench()
    rng = np.random.default_rng(9)
    acc = _acc(rng, 64, lambda r, sh: np.ones(sh))
    ntau, beta = 8, 2.0
    per_block = acc["gtau"] / acc["scalars"][:, 0][:, None, None] * ntau / beta ** 2
    ref = b.gtau_inv_err2({"gtau": per_block})
    jk = b.ratio_inv_err2(acc["gtau"], acc["scalars"][:, 0], ntau, beta)
    assert np.isclose(jk, ref, rtol=1e-10)


def test_group_ratio_estimator_finite_with_zero_sign_chains():
    b = _bench()
    rng = np.random.default_rng(8)
    acc = _acc(rng, 256, lambda r, sh: np.where(r.random(sh) < 0.8, 1.0, -1.0))
    acc["scalars"][:3, 0] = 0.0        # chains whose own sign sum vanished
    g, s = _groups(acc, 32)
    v = b.ratio_inv_err2(g, s, 8, 2.0)
    assert np.isfinite(v) and v > 0


def test_binning_plateau_recovers_autocorrelated_error():
    """An AR(1) series x_t = rho x_{t-1} + sqrt(1 - rho^2) e_t of unit variance
    has var(mean) -> (1 + rho) / (1 - rho) / n; the plateau of the binning
    accumulator (the oracle's autocorracc) recovers it, the unbinned level
    underestimates it by that factor."""
    import oracle as O
    b = _bench()
    rng = np.random.default_rng(11)
    n, rho, nlev = 1 << 15, 0.8, 15
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0]
    for t in range(1, n):
        x[t] = rho * x[t - 1] + np.sqrt(1 - rho ** 2) * e[t]
    _, blocks = O.autocorr(x[:, None], nlev)
    # ntau / beta^2 = 1 and sign 1: the plateau is the error of the mean itself
    inv, curve = b.plateau_inv_err2(blocks, n, 1.0, 1, 1.0, min_bins=64)
    exact = (1 + rho) / (1 - rho) / n
    assert abs(1.0 / inv / exact - 1.0) < 0.35, (1.0 / inv, exact)
    assert curve[0] < 0.3 * exact
