"""CPU tests of the oracle: pinned against the reference's golden vectors
(tests/golden/*.npz, written by tests/golden/make_golden.py from the reference
core built from its own sources) and against closed forms.
"""

import math

import numpy as np
import pytest

from conftest import load_golden, qmc_of
from oracle import OracleChain, OracleModel, rng_stream
from w2dynamics_amd import problems as P

LOG_CASES = ["c1_log_w1", "c1_log_win", "c2_log_win", "c3_log_win", "c4_log_w1", "c3_log_all", "c3d_log_all", "c4_log_all",
             "c2_log_sym", "c3_log_sym"]
FIELDS = ("kind", "outcome", "noper", "outer_sst", "window")


# ---------------------------------------------------------------- RNG -------
def test_mt19937_known_answer():
    """C++ [rand.predef]: the 10000th output of a default-constructed mt19937
    (seed 5489) is 4123659995."""
    out = rng_stream(5489, np.full(10000, 2))
    assert int(out[-1]) == 4123659995


def test_grnd_is_53bit_canonical():
    """generate_canonical<double,53> from two 32-bit draws: (g1 + g2 2^32) / 2^64."""
    raw = rng_stream(42, np.full(2000, 2))
    d = rng_stream(42, np.zeros(1000, np.int32))
    g1, g2 = raw[0::2], raw[1::2]
    expect = (g1 + g2 * 2.0 ** 32) / 2.0 ** 64
    assert np.array_equal(d, expect)
    assert (d >= 0).all() and (d < 1).all()


def test_randint_lemire_range_and_uniformity():
    n = 20000
    lo = np.full(n, 1, np.int32)
    hi = np.full(n, 6, np.int32)
    v = rng_stream(7, np.ones(n, np.int32), lo, hi)
    assert v.min() == 1 and v.max() == 6
    counts = np.bincount(v.astype(int))[1:]
    assert np.all(np.abs(counts - n / 6) < 5 * math.sqrt(n / 6))
    # range 1 consumes exactly one draw and returns the lower bound
    w = rng_stream(7, np.array([1, 2]), np.array([3, 0]), np.array([3, 0]))
    assert w[0] == 3 and w[1] == rng_stream(7, np.array([2, 2]))[1]


# ------------------------------------------------------------- model --------
@pytest.mark.parametrize("name", ["c1_log_w1", "c2_log_win", "c3_log_win", "c4_log_w1", "c3_log_all", "c3d_log_all",
                                  "c4_log_all"])
def test_model_matches_reference(name):
    z = load_golden(name)
    m = OracleModel(z["umatrix"], z["muimp"], qmc_of(z)["QuantumNumbers"])
    assert m.nsst == int(z["model_nsst"])
    assert np.array_equal(m.sizes, z["model_sizes"])
    assert np.array_equal(np.concatenate(m.states), z["model_states"])
    assert np.array_equal(m.connect, z["model_connect"])
    ev = np.concatenate(m.evals)
    assert np.abs(ev - z["model_evals"]).max() <= 1e-11 * max(1.0, np.abs(ev).max())
    assert abs(m.egs - float(z["model_egs"])) <= 1e-11 * max(1.0, abs(m.egs))


def test_model_spectrum_matches_fock_space_ed():
    """Block spectra equal the eigenvalues of the full Fock-space Hamiltonian."""
    p = P.kanamori_3orb(beta=20.0)
    m = OracleModel(p.umatrix, p.muimp)
    h = P.fock_hamiltonian(p.umatrix, p.muimp)
    assert np.allclose(np.sort(np.concatenate(m.evals)), np.linalg.eigvalsh(h), atol=1e-10)


def test_eigenbasis_operators_anticommute():
    """{c_i, c+_j} = delta_ij on every block (eigenbasis blocks)."""
    p = P.kanamori_2orb(beta=10.0)
    m = OracleModel(p.umatrix, p.muimp)
    nb = m.nbands
    for s in range(m.nsst):
        n = m.sizes[s]
        for b1 in range(nb):
            for s1 in range(2):
                for b2 in range(nb):
                    for s2 in range(2):
                        acc = np.zeros((n, n))
                        t = m.connect[s, b2, s2, 0]          # c+_2 first
                        if t >= 0 and m.connect[t, b1, s1, 1] == s:
                            acc += m.psi_eb[(t, b1, s1, 1)] @ m.psi_eb[(s, b2, s2, 0)]
                        t = m.connect[s, b1, s1, 1]          # c_1 first
                        if t >= 0 and m.connect[t, b2, s2, 0] == s:
                            acc += m.psi_eb[(t, b2, s2, 0)] @ m.psi_eb[(s, b1, s1, 1)]
                        want = np.eye(n) if (b1, s1) == (b2, s2) else np.zeros((n, n))
                        assert np.allclose(acc, want, atol=1e-12)


# --------------------------------------------------------- move logs --------
@pytest.mark.parametrize("name", LOG_CASES)
def test_oracle_move_log_matches_reference(name):
    z = load_golden(name)
    q = qmc_of(z)
    ch = OracleChain(z["umatrix"], z["ftau"], z["muimp"], float(z["beta"]), q, int(z["seed"]))
    ch.set_taudiffmax(float(q["TaudiffMax"]))
    rec = ch.moves(int(z["nmoves"]))
    for f in FIELDS:
        bad = np.nonzero(rec[f] != z["move_" + f])[0]
        assert len(bad) == 0, "%s: %s differs first at move %d" % (name, f, bad[0])
    assert np.abs(rec["trace_log"] - z["move_trace_log"]).max() < 1e-10
    assert np.abs(rec["det_log"] - z["move_det_log"]).max() < 1e-10
    for f in ("trace_sign", "det_sign", "cfgsign"):
        assert np.array_equal(rec[f], z["move_" + f])
    # final configuration and inverse hybridisation blocks
    cfg = ch.config()
    assert cfg["n"] == int(z["cfg_n"])
    assert np.allclose(cfg["taus"], z["cfg_taus"], rtol=0, atol=0)
    nf = 2 * int(z["nbands"])
    for f in range(nf):
        a, b = ch.minv(f), z["cfg_minv_%d" % f]
        assert a.shape == b.shape
        if a.size:
            assert np.abs(a - b).max() <= 1e-9 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("name", ["c1_run", "c2_run"])
def test_oracle_observables_match_reference(name):
    z = load_golden(name)
    q = qmc_of(z)
    ch = OracleChain(z["umatrix"], z["ftau"], z["muimp"], float(z["beta"]), q, int(z["seed"]))
    ch.set_taudiffmax(float(q["TaudiffMax"]))
    ch.warmup(int(z["nwarm"]))
    ch.measure(int(q["Nmeas"]), int(q["NCorr"]))
    r = ch.results()
    for k in ("mean_sign", "gtau", "gleg", "occ", "densitymatrix", "histo", "acc",
              "outer_sst_histo"):
        a, b = np.asarray(r[k]), np.asarray(z["res_" + k])
        assert a.shape == b.shape, k
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12), (k, np.abs(a - b).max())


# ------------------------------------------------------ closed forms --------
def test_empty_trace_is_boltzmann_sum():
    p = P.kanamori_2orb(beta=10.0)
    ch = OracleChain.from_problem(p, 1)
    m = ch.model
    for s in range(m.nsst):
        w = ch.set_config([], [], [], [], s)
        expect = np.sum(np.exp(-p.beta * (m.evals[s] - m.egs)))
        assert w["trace_sign"] == 1.0
        assert abs(w["trace_log"] - math.log(expect)) < 1e-12


def test_two_operator_trace_closed_form():
    """Tr_s[e^{-(beta-t2) H} c e^{-(t2-t1) H} c+ e^{-t1 H}] from the eigenbasis."""
    p = P.kanamori_2orb(beta=8.0)
    ch = OracleChain.from_problem(p, 3)
    m = ch.model
    t1, t2 = 1.3, 5.2
    for s in range(m.nsst):
        for b in range(2):
            for sp in range(2):
                t = m.connect[s, b, sp, 0]
                if t < 0 or m.connect[t, b, sp, 1] != s:
                    continue
                cd = m.psi_eb[(s, b, sp, 0)]
                c = m.psi_eb[(t, b, sp, 1)]
                es, et = m.evals[s] - m.egs, m.evals[t] - m.egs
                mat = (np.diag(np.exp(-(p.beta - t2) * es)) @ c @ np.diag(np.exp(-(t2 - t1) * et))
                       @ cd @ np.diag(np.exp(-t1 * es)))
                val = np.trace(mat)
                w = ch.set_config([t1, t2], [b, b], [sp, sp], [1, 2], s)
                assert abs(w["trace_sign"] * math.exp(w["trace_log"]) - val) <= 1e-12 * abs(val)


def test_configuration_breaking_quantum_numbers_has_zero_weight():
    """An annihilator acting first on the empty superstate has no Connect target:
    the walk stops with a zero trace instead of reading past the block tables,
    and the same pair on a superstate that has the electron stays finite."""
    p = P.kanamori_2orb(beta=8.0)
    ch = OracleChain.from_problem(p, 3)
    m = ch.model
    empty = next(s for s in range(m.nsst) if 0 in m.states[s])
    w = ch.set_config([1.0, 2.0], [0, 0], [0, 0], [2, 1], empty)
    assert w["trace_sign"] == 0.0 and w["trace_log"] < -1e300
    full = next(s for s in range(m.nsst) if m.connect[s, 0, 0, 1] >= 0)
    w = ch.set_config([1.0, 2.0], [0, 0], [0, 0], [2, 1], full)
    assert w["trace_sign"] != 0.0 and np.isfinite(w["trace_log"])


def test_determinant_and_inverse_brute_force():
    """det and M = Delta^{-1} of a visited configuration against numpy."""
    p = P.kanamori_2orb(beta=20.0, TaudiffMax=4.0)
    ch = OracleChain.from_problem(p, 5)
    ch.set_taudiffmax(4.0)
    ch.moves(4000, record=False)
    cfg = ch.config()
    nb = 2
    beta = p.beta
    f_of = lambda b, s: b + nb * s  # noqa: E731
    ftab = p.ftau
    nft = ftab.shape[-1]
    tot = 0.0
    sgn = 1.0
    for f in range(2 * nb):
        b, s = f % nb, f // nb
        sel = (cfg["orbs"] == b) & (cfg["spins"] == s)
        tc = cfg["taus"][sel & (cfg["cas"] == 1)]
        ta = cfg["taus"][sel & (cfg["cas"] == 2)]
        if len(tc) == 0:
            continue

        def F(t):
            sig = np.where(t < 0, -1.0, 1.0)
            t = np.where(t < 0, t + beta, t)
            n = t / beta * (nft - 1)
            i = n.astype(int)
            g = ftab[b, s, b, s]
            return sig * (g[i] + (n - i) * (g[i + 1] - g[i]))
        D = F(ta[:, None] - tc[None, :])
        sd, ld = np.linalg.slogdet(D)
        tot += ld
        sgn *= sd
        assert np.allclose(ch.minv(f), np.linalg.inv(D), rtol=1e-8, atol=1e-10)
        assert np.allclose(ch.minv(f) @ D, np.eye(len(tc)), atol=1e-9)
    assert abs(cfg["det_log"] - tot) < 1e-9
    assert cfg["det_sign"] == sgn
    del f_of


def test_legendre_recursion_matches_numpy():
    """The Bonnet recursion of MeasGtauRem equals numpy's Legendre polynomials."""
    from numpy.polynomial import legendre as L
    x = np.linspace(-1, 1, 37)
    pl = np.zeros_like(x)
    plm1 = np.zeros_like(x)
    for l in range(30):
        if l == 0:
            plp1 = np.ones_like(x)
        elif l == 1:
            plp1 = x.copy()
        else:
            ll = l - 1
            plp1 = (2 * ll + 1) / (ll + 1) * x * pl - ll / (ll + 1) * plm1
        assert np.allclose(plp1, L.legval(x, np.eye(30)[l]), atol=1e-12)
        plm1, pl = pl, plp1


def legendre_to_matsubara(nleg: int, niw: int) -> np.ndarray:
    """T[n, l] = (-1)^n i^(l+1) sqrt(2l+1) j_l((2n+1) pi / 2) for n = -niw..niw-1
    (Boehnke et al. 2011, eq. E2; the transform w2dyn/auxiliaries/transform.py
    leg2mat documents)."""
    from scipy.special import spherical_jn
    n = np.arange(-niw, niw)
    l = np.arange(nleg)
    z = (2 * n + 1) * np.pi / 2
    # j_l(-x) = (-1)^l j_l(x)
    jl = spherical_jn(l[None, :], np.abs(z)[:, None]) * np.sign(z)[:, None] ** l[None, :]
    return ((-1.0) ** n)[:, None] * (1j ** (l + 1))[None, :] * np.sqrt(2 * l + 1)[None, :] * jl


def test_giw_is_the_matsubara_transform_of_the_legendre_estimator():
    """Every pair term of G(i w_n) is exp(i w_n dtau) and of G_l is P_l(x) on the
    same M: by the plane-wave (Rayleigh) expansion exp(i z x) = sum_l (2l+1)
    i^l j_l(z) P_l(x), G(i w_n) = sum_l T_nl G_l exactly once NLegMax exceeds
    w_n beta / 2 by a margin.  A sign, conjugation, transposed M or shifted
    frequency fails this."""
    niw, nleg = 16, 120
    p = P.kanamori_2orb(beta=10.0, MeasGiw=1, Niw=niw, NLegMax=nleg)
    ch = OracleChain.from_problem(p, 7)
    ch.warmup(2000)
    ch.measure(300, 50)
    r = ch.results()
    giw = r["giw"]
    assert giw.shape == (2, 2, 2 * niw)
    want = np.einsum("nl,bsl->bsn", legendre_to_matsubara(nleg, niw), r["gleg"])
    assert np.abs(giw - want).max() <= 1e-12 * np.abs(giw).max()
    # real estimator: G(i w_{-n-1}) = conj G(i w_n)
    assert np.array_equal(giw[..., :niw], np.conj(giw[..., niw:][..., ::-1]))


def test_giw_frequency_grid_follows_the_nfft_fallback():
    """One measurement of a fixed configuration against get_giw
    (Nfft_z.F90:133-203) computed with the fallback adjoint transform
    (Nfft_z.F90:64-90): 2 nfreq = 4 Niw frequencies on the grid
    pi (2n - nn_total - 2), the shift exp(+-i pi tau / beta), every other output;
    times -cfgsign (MeasGtauRem's lookup term) and zScaling / beta."""
    niw, beta = 6, 20.0
    p = P.kanamori_2orb(beta=beta, MeasGiw=1, Niw=niw)
    ch = OracleChain.from_problem(p, 11)
    ch.warmup(5000)
    cfg = ch.config()
    assert min(np.sum((cfg["orbs"] == b) & (cfg["spins"] == s)) for b in range(2) for s in range(2)) >= 4
    ch.measure(1, 0)   # no moves: one measurement of this configuration
    giw = ch.results()["giw"]
    nfreq = 2 * niw
    nn = 2 * nfreq
    grid = np.pi * (2 * np.arange(1, nn + 1) - nn - 2)
    cfgsign = cfg["cfgsign"]
    for b in range(2):
        for s in range(2):
            f = b + 2 * s
            sel = [(t, c) for t, o, sp, c in zip(cfg["taus"], cfg["orbs"], cfg["spins"], cfg["cas"])
                   if o == b and sp == s]
            tc = np.array([t for t, c in sel if c == 1])
            ta = np.array([t for t, c in sel if c == 2])
            m = ch.minv(f)   # rows creators, columns annihilators, time order
            exp_c = np.exp(-1j * tc * np.pi / beta)
            exp_a = np.exp(1j * ta * np.pi / beta)
            func = (exp_c[None, :] * m.T * exp_a[:, None]).ravel()          # func(i, j)
            x = ((ta[:, None] - tc[None, :]) / (2 * beta)).ravel()          # dtau(i, j)
            fhat = (np.exp(1j * grid[:, None] * x[None, :]) * func[None, :]).sum(1)
            ref = -cfgsign * fhat[0::2] * (-1.0 / (1 * cfgsign)) / beta
            assert np.abs(giw[b, s] - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300)


def _anderson_ed(eps_d, u, eps_k, v_k, beta):
    """Exact <n_up> and <n_up n_dn> of a single-orbital Anderson impurity with a
    discrete bath, from the full Fock space (bits: d_up, d_dn, then bath)."""
    nbath = len(eps_k)
    nso = 2 + 2 * nbath
    dim = 1 << nso

    def mode(site, s):   # site 0 = impurity, 1.. = bath
        return 2 * site + s

    def hop(state, a, b):   # c+_a c_b |state>
        if not (state >> b) & 1:
            return None, 0
        sg = (-1) ** bin(state & ((1 << b) - 1)).count("1")
        st = state ^ (1 << b)
        if (st >> a) & 1:
            return None, 0
        sg *= (-1) ** bin(st & ((1 << a) - 1)).count("1")
        return st ^ (1 << a), sg

    h = np.zeros((dim, dim))
    for st in range(dim):
        nu, nd = (st >> 0) & 1, (st >> 1) & 1
        e = eps_d * (nu + nd) + u * nu * nd
        for k in range(nbath):
            for s in range(2):
                e += eps_k[k] * ((st >> mode(k + 1, s)) & 1)
        h[st, st] = e
        for k in range(nbath):
            for s in range(2):
                for a, b in ((mode(0, s), mode(k + 1, s)), (mode(k + 1, s), mode(0, s))):
                    t, sg = hop(st, a, b)
                    if t is not None:
                        h[t, st] += v_k[k] * sg
    w, vec = np.linalg.eigh(h)
    p = np.exp(-beta * (w - w[0]))
    p /= p.sum()
    nu = np.array([(st >> 0) & 1 for st in range(dim)], float)
    nd = np.array([(st >> 1) & 1 for st in range(dim)], float)
    prob = (vec ** 2) @ p
    return float(prob @ nu), float(prob @ (nu * nd))


def test_full_run_matches_exact_diagonalisation():
    """Impurity + 3 discrete bath sites: <n_up> and double occupancy of the
    sampled density matrix agree with exact diagonalisation within 4 sigma
    (sigma from independent chains)."""
    beta, u, mu = 5.0, 2.0, 0.6
    eps_k, v_k = np.array([-1.0, 0.2, 1.1]), np.array([0.5, 0.35, 0.45])
    p = P.anderson_1orb(beta=beta, u=u, nftau=4001)
    p.ftau = P.discrete_bath_ftau(beta, eps_k, v_k, 4001)
    p.muimp = P.solver_muimp(1, mu)
    n_ed, d_ed = _anderson_ed(-mu, u, eps_k, v_k, beta)
    nup, docc = [], []
    for seed in range(8):
        ch = OracleChain.from_problem(p, 100 + seed)
        ch.warmup(5000)
        ch.measure(2000, 20)
        r = ch.results()
        nup.append(r["occ"][0, 0, 0, 0])
        docc.append(r["occ"][0, 0, 0, 1])
    for vals, exact in ((nup, n_ed), (docc, d_ed)):
        m = np.mean(vals)
        se = np.std(vals, ddof=1) / np.sqrt(len(vals))
        assert abs(m - exact) < 4 * se + 2e-3, (m, se, exact)


def test_binning_accumulator_closed_form():
    """autocorracc (Accumulators.F90:283-362) on the sequence 0, 1, ..., 2^n - 1
    (the reference's own test program, Accumulators.F90:366-410): the bin
    means of level l form an arithmetic sequence of step 2^l and length
    N = 2^(n-l), whose variance is 4^l N (N + 1) / 12; the mean is (2^n - 1)/2."""
    import oracle as O
    n = 12
    data = np.arange(2 ** n, dtype=float)[:, None]
    mean, blocks = O.autocorr(data, n)
    assert abs(mean[0] - (2 ** n - 1) / 2) < 1e-9
    for lvl in range(n):
        N = 2 ** (n - lvl)
        exact = 4.0 ** lvl * N * (N + 1) / 12.0
        assert abs(blocks[lvl, 0] - exact) <= 1e-9 * exact, (lvl, blocks[lvl, 0], exact)


def test_binning_accumulator_matches_bin_means():
    """On random data of a length that is not a power of two, level l holds the
    sample variance of the means of the complete 2^l-bins."""
    import oracle as O
    rng = np.random.default_rng(3)
    n, nc, nlev = 1000, 3, 6
    data = rng.normal(size=(n, nc)) + np.arange(nc)
    mean, blocks = O.autocorr(data, nlev)
    assert np.allclose(mean, data.mean(0), rtol=0, atol=1e-12)
    for lvl in range(nlev):
        w = 2 ** lvl
        nb = n // w
        bins = data[:nb * w].reshape(nb, w, nc).mean(1)
        ref = bins.var(0, ddof=1) * (nb - 1) / ((n >> lvl) - 1)
        assert np.allclose(blocks[lvl], ref, rtol=1e-10), lvl


def _calib_flow(chain, z, q):
    """cthyb_solver.py:899-944 for one rank up to the phase-2 histogram, then
    warm-up and measurement; returns the phase-1 list, TaudiffMax and the
    phase-2 AccPair."""
    chain.init_counters()
    chain.calibrate(True, 1000, 100, 1000, 100000)
    lst = np.asarray(chain.pair_stats()["accpairtau"] if hasattr(chain, "pair_stats") else [])
    tdm = float(np.sort(lst)[int(0.995 * 1000)])
    chain.set_taudiffmax(tdm)
    chain.init_counters()
    chain.calibrate(False, 1000, 100, 1000, 100000)
    ap = chain.pair_stats()["accpair"]
    return lst, tdm, ap


def _blocks_of(z):
    sizes = np.asarray(z["model_sizes"])
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    return [(int(o), int(o + n)) for o, n in zip(offs, sizes)]


def calib_raw_oracle(z):
    """The calibration fixture's flow on the oracle: raw arrays for the host
    normalisation."""
    q = qmc_of(z)
    o = OracleChain(z["umatrix"], z["ftau"], z["muimp"], float(z["beta"]), q, int(z["seed"]),
                    q["QuantumNumbers"])
    lst, tdm, ap2 = _calib_flow(o, z, q)
    o.warmup(int(z["nwarm"]))
    o.measure(int(z["nmeas"]), int(q["NCorr"]))
    r = o.raw()
    h = o.outer_histos()
    st = o.steps()
    acc = {k: v[None] for k, v in h.items()}
    acc.update(outer_hist=r["outer_hist"][None], accpair=o.pair_stats()["accpair"][None],
               gtau_mean_step=st["gtau_mean_step"][None], gtau_mid_step=st["gtau_mid_step"][None],
               sign_step=st["sign_step"][None])
    return lst, tdm, ap2, acc, r["scalars"][1], o.results()


def test_oracle_calibration_matches_reference():
    """ctqmc_calibrate phases 1 and 2 (AccPairTau, the TaudiffMax of its 99.5 %
    point, AccPair over 100 bins of [0, TaudiffMax]) and the run after it:
    AccPair, outer-state sign / visit / trace-contribution histograms and the
    step-resolved G(tau) and sign (CTQMC.F90:2394-2402, 3285-3339, 8741-8764),
    normalised by the host's post-processing, against the reference."""
    from w2dynamics_amd.ctqmc import normalise_histograms
    z = load_golden("c2_calib")
    lst, tdm, ap2, acc, cnt, res = calib_raw_oracle(z)
    assert np.allclose(lst, z["calib_accpairtau"], rtol=1e-12, atol=1e-12)
    assert tdm == float(z["calib_taudiffmax"])
    nb = int(z["nbands"])
    assert np.array_equal(ap2.reshape(2, nb, -1).swapaxes(0, 1), z["calib_accpair"])
    h = normalise_histograms(acc, np.array([cnt]), nb, float(z["beta"]))
    assert np.array_equal(h["accpair"][0], z["res_accpair"])
    for k in ("sign_sst", "contrib_sst"):
        assert np.allclose(h[k][0], z["res_" + k], rtol=1e-10, atol=1e-14, equal_nan=True), k
    # per eigenstate: the order of degenerate eigenstates within a block is the
    # eigensolver's choice, so compare the values of each block as sets
    blocks = _blocks_of(z)
    for k, ref in (("sign_state", "res_sign_state"), ("contrib_state", "res_contrib_state"),
                   ("outer_state_histo", "res_outer_state")):
        for a, b in blocks:
            assert np.allclose(np.sort(h[k][0][a:b]), np.sort(z[ref][a:b]), rtol=1e-10, atol=1e-14,
                               equal_nan=True), (k, a, b)
    for k in ("gtau_mean_step", "gtau_mid_step"):
        assert np.allclose(h[k][0], z["res_" + k], rtol=1e-10, atol=1e-14), k
    assert np.array_equal(h["sign_step"][0], z["res_sign_step"])
    assert np.allclose(res["gtau"], z["res_gtau"], rtol=1e-10, atol=1e-13)
