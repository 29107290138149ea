"""CPU tests of the C-ABI library and the host-side logic (no GPU compute)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from w2dynamics_amd import ctqmc
from w2dynamics_amd import problems as P

HEADER = os.path.join(ROOT, "include", "w2dyn_ctqmc.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(w2d_ctqmc_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctqmc.load_library()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), "missing export " + n
    assert b"gfx950" in lib.w2d_ctqmc_version()


def test_library_missing_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        ctqmc._lib = None
        try:
            ctqmc.load_library(str(tmp_path / "nope.so"))
        finally:
            ctqmc._lib = None
    ctqmc.load_library()


@pytest.mark.parametrize("override", [{"offdiag": 1}, {"segment": 1}, {"statesampling": 1},
                                      {"Percentage4OperatorMove": 0.1},
                                      {"PercentageOuterMove": 0.1},
                                      {"PercentageWormInsert": 0.2}])
def test_unsupported_branches_are_rejected(override):
    """Parameter checks happen before any HIP call and return W2D_ERR_ARG."""
    p = P.anderson_1orb(**override)
    with pytest.raises(ctqmc.CtqmcError) as e:
        ctqmc.CtqmcGpu.from_problem(p)
    assert e.value.code == 1


def test_null_arguments_are_rejected():
    lib = ctqmc.load_library()
    h = C.c_void_p()
    assert lib.w2d_ctqmc_create(None, 1, 10, None, None, None, 1, 0, 0, 0, C.byref(h)) == 1
    assert lib.w2d_ctqmc_warmup(None, 10) == 1
    assert lib.w2d_ctqmc_destroy(None) == 0


def test_param_string_roundtrip():
    p = P.kanamori_3orb(Nmeas=123)
    s = p.param_string()
    assert ctqmc._param_value(s, "beta") == p.beta
    assert "QuantumNumbers = Nt Szt #" in s
    assert ctqmc._param_value(s, "Nmeas") == 123


def test_normalise_matches_reference_formulas():
    """CTQMC.F90:8663-8771 on synthetic accumulators of two chains."""
    nb, ntau, nleg, beta = 1, 4, 3, 2.0
    acc = dict(gtau=np.arange(2 * 2 * ntau, dtype=float).reshape(2, 2, ntau),
               gleg=np.ones((2, 2, nleg)), dm=np.zeros((2, 4, 4)),
               histo=np.zeros((2, 2, 6)), outer_hist=np.ones((2, 4)),
               scalars=np.array([[-10.0, 10.0] + [1.0] * 12, [-5.0, 10.0] + [1.0] * 12]))
    model = dict(offsets=np.arange(4), sizes=np.ones(4, int), evecs=[np.eye(1)] * 4,
                 states=np.array([0, 1, 2, 3]))
    r = ctqmc.normalise(acc, nb, beta, ntau, nleg, model)
    assert np.allclose(r["mean_sign"], [-1.0, -0.5])
    g0 = acc["gtau"][0].reshape(2, 1, ntau).swapaxes(0, 1) / 10.0 / -1.0 * ntau / beta ** 2
    assert np.allclose(r["gtau"][0], g0)
    zs = -1.0 / (10.0 * -0.5)
    gl = -np.ones((1, 2, nleg)) * np.sqrt(2 * np.arange(nleg) + 1) / beta * zs
    assert np.allclose(r["gleg"][1], gl)


def test_coulomb_u_reduces_to_density_for_f0_only():
    """With only F0 the Coulomb U is U_{abab} = F0 for all orbital pairs."""
    u = P.u_coulomb(5, 4.0)
    for a in range(5):
        for b in range(5):
            assert abs(u[a, 0, b, 1, a, 0, b, 1] - 4.0) < 1e-12
    assert abs(u[0, 0, 1, 1, 1, 0, 0, 1]) < 1e-12   # no exchange without F2/F4


def test_coulomb_u_hermitian_symmetries():
    u = P.u_coulomb(5, *P.slater_from_uj_d(4.0, 0.8))
    n = 10
    uf = u.reshape(n, n, n, n)
    assert np.allclose(uf, uf.transpose(1, 0, 3, 2))          # particle exchange
    assert np.allclose(uf, uf.transpose(2, 3, 0, 1))          # hermiticity (real U)


def test_semicircle_ftau_sum_rule():
    """F(0) + F(beta) = V^2 = D^2/4 and F is positive and symmetric at half filling."""
    f = P.semicircle_ftau(1, 10.0, 2.0, 501)[0, 0, 0, 0]
    assert abs(f[0] + f[-1] - 1.0) < 1e-10
    assert np.allclose(f, f[::-1], atol=1e-12)
    assert (f > 0).all()


def _synthetic_sums(signs, cnt=10.0, nb=1, ntau=4, nleg=3, seed=0):
    rng = np.random.default_rng(seed)
    n = len(signs)
    sc = np.zeros((n, 14))
    sc[:, 0] = signs
    sc[:, 1] = cnt
    sc[:, 2:12] = 1.0
    return dict(gtau=rng.normal(size=(n, 2 * nb, ntau)), gleg=rng.normal(size=(n, 2 * nb, nleg)),
                dm=rng.normal(size=(n, 4, 4)), histo=np.ones((n, 2 * nb, 6)),
                outer_hist=np.ones((n, 4)), scalars=sc)


MODEL1 = dict(offsets=np.arange(4), sizes=np.ones(4, int), evecs=[np.eye(1)] * 4,
              states=np.array([0, 1, 2, 3]))


def test_pooled_estimate_is_ratio_of_sums():
    """The pooled estimate is sum(G) / sum(sign) (CTQMC.F90:8663-8771 on pooled
    sums), and its jackknife error is the textbook leave-one-out formula."""
    nb, ntau, nleg, beta = 1, 4, 3, 2.0
    s = _synthetic_sums([-10.0, -6.0, -8.0, -9.0])
    r = ctqmc.pooled_estimate(s, nb, beta, ntau, nleg, MODEL1)
    g = s["gtau"].sum(0).reshape(2, 1, ntau).swapaxes(0, 1) / s["scalars"][:, 0].sum() * ntau / beta ** 2
    assert np.allclose(r["gtau"], g)
    assert np.isclose(r["mean_sign"], s["scalars"][:, 0].sum() / 40.0)
    jk = np.array([(s["gtau"].sum(0) - s["gtau"][i]) / (s["scalars"][:, 0].sum() - s["scalars"][i, 0])
                   for i in range(4)]) * ntau / beta ** 2
    err = np.sqrt(3 / 4 * ((jk - jk.mean(0)) ** 2).sum(0))
    assert np.allclose(r["gtau_err"], err.reshape(2, 1, ntau).swapaxes(0, 1))


def test_pooled_giw_is_ratio_of_sums_with_complex_jackknife():
    """G(i w_n) pooled: sum(raw) / (sum(sign) beta) for n >= 0, the conjugates for
    n < 0 (Giw * zScaling / beta, CTQMC.F90:8816, on pooled sums), and a
    jackknife error over the real and the imaginary part together."""
    nb, ntau, nleg, beta, niw = 1, 4, 3, 2.0, 5
    s = _synthetic_sums([-10.0, -6.0, -8.0, -9.0])
    s["giw"] = np.random.default_rng(3).normal(size=(4, 2 * nb, niw, 2))
    r = ctqmc.pooled_estimate(s, nb, beta, ntau, nleg, MODEL1)
    raw = s["giw"][..., 0] + 1j * s["giw"][..., 1]
    pos = raw.sum(0) / (s["scalars"][:, 0].sum() * beta)
    assert r["giw"].shape == (nb, 2, 2 * niw)
    assert np.allclose(r["giw"][..., niw:], pos.reshape(2, nb, niw).swapaxes(0, 1))
    assert np.array_equal(r["giw"][..., :niw], np.conj(r["giw"][..., niw:][..., ::-1]))
    jk = np.array([(raw.sum(0) - raw[i]) / ((s["scalars"][:, 0].sum() - s["scalars"][i, 0]) * beta)
                   for i in range(4)])
    err = np.sqrt(3 / 4 * (np.abs(jk - jk.mean(0)) ** 2).sum(0))
    assert np.allclose(r["giw_err"][..., niw:], err.reshape(2, nb, niw).swapaxes(0, 1))


def test_pooled_estimate_survives_a_zero_sign_chain():
    """One chain whose sign sum vanishes: its own normalisation is undefined
    (NaN, as a reference rank would divide by zero) but the pooled estimate and
    its error stay finite."""
    nb, ntau, nleg, beta = 1, 4, 3, 2.0
    s = _synthetic_sums([-10.0, 0.0, -8.0, -9.0])
    per_chain = ctqmc.normalise(s, nb, beta, ntau, nleg, MODEL1)
    assert not np.all(np.isfinite(per_chain["gtau"][1]))
    r = ctqmc.pooled_estimate(s, nb, beta, ntau, nleg, MODEL1)
    for k in ("gtau", "gleg", "densitymatrix", "occ", "mean_sign"):
        assert np.all(np.isfinite(r[k])) and np.all(np.isfinite(r[k + "_err"])), k


def test_reduced_density_matrices_match_reference():
    """rho1 and rho2 from the reference's occupation-basis density matrix agree
    with the reference's own Meas_rho1/rho2_from_Densitymatrix (CTQMC.F90:2936-3060)."""
    from conftest import load_golden
    z = load_golden("c2_calib")
    r1, r2 = ctqmc.reduced_density_matrices(np.asarray(z["res_densitymatrix"]), int(z["nbands"]),
                                            np.asarray(z["model_states"]))
    assert np.allclose(r1, z["res_rho1"], rtol=1e-10, atol=1e-13)
    assert np.allclose(r2, z["res_rho2"], rtol=1e-10, atol=1e-13)


def test_accumulation_guard_catches_a_lost_bin():
    """The G(tau) bin sum equals the l=0 Legendre sum on consistent
    accumulators (here the oracle's) and the guard raises when one bin's
    contribution is lost."""
    import oracle as O
    prob = P.kanamori_2orb(beta=10.0, TaudiffMax=3.0, Ntau=40, NLegMax=8)
    o = O.OracleChain.from_problem(prob, 3)
    o.set_taudiffmax(3.0)
    o.warmup(2000)
    o.measure(50, 20)
    r = o.raw()
    acc = {"gtau": r["gtau"][None], "gleg": r["gleg"][None]}
    ctqmc.check_accumulators(acc)
    i = int(np.argmax(np.abs(r["gtau"][0])))
    acc["gtau"] = acc["gtau"].copy()
    acc["gtau"][0, 0, i] = 0.0              # one bin's atomic adds lost
    with pytest.raises(ctqmc.AccumulationError):
        ctqmc.check_accumulators(acc)
