"""Impurity-problem inputs for the CT-HYB sampling core.

These builders produce exactly the arrays the reference solver module receives
through ``init_solver`` (``src/ctqmc_fortran/CTQMC.F90:7894``):

* ``umatrix``  -- (N,2,N,2,N,2,N,2) interaction ``U_{1234}`` in the convention
  ``H_U = 1/2 sum U_{1234} c+_1 c+_2 c_4 c_3`` (``Operator.F90:1066-1095``);
  builders follow ``w2dyn/dmft/interaction.py:171-433``.
* ``ftau``     -- (N,2,N,2,Nftau) hybridisation ``F(tau)`` on an equidistant
  grid including both ends, diagonal in flavour, positive for ``0<tau<beta``
  (the hole-propagator convention of ``cthyb:190-198``); for a bath of discrete
  sites it is ``sum |V|^2/(exp(-tau e)+exp((beta-tau) e))``
  (``w2dyn/auxiliaries/transform.py:114-146``).
* ``muimp``    -- (N,2,N,2) one-particle term as the solver sees it, i.e. the
  *negated* ``mu*1 - crystalfield`` (``cthyb:160-164``,
  ``w2dyn/dmft/cthyb_solver.py:255``).

plus the ``key = value #`` parameter string parsed by
``Parameters.F90:281-320`` (built like ``cthyb_solver.py:764-774``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------
# interaction matrices
# --------------------------------------------------------------------------


def u_density(nbands: int, u: float, v: float, j: float) -> np.ndarray:
    """Density-density U (``w2dyn/dmft/interaction.py:171-192``)."""
    um = np.zeros((nbands, 2) * 4)
    for b1 in range(nbands):
        for s in (0, 1):
            um[b1, s, b1, 1 - s, b1, s, b1, 1 - s] = u
        for b3 in range(nbands):
            if b3 == b1:
                continue
            for s1 in (0, 1):
                for s3 in (0, 1):
                    um[b1, s1, b3, s3, b1, s1, b3, s3] = v
            for s in (0, 1):
                um[b1, s, b3, s, b3, s, b1, s] = j
    return um


def u_kanamori(nbands: int, u: float, v: float, j: float) -> np.ndarray:
    """Kanamori U: density part plus pair hopping and spin flip
    (``w2dyn/dmft/interaction.py:221-244``)."""
    um = u_density(nbands, u, v, j)
    for b1 in range(nbands):
        for b2 in range(nbands):
            if b1 == b2:
                continue
            for s in (0, 1):
                um[b1, s, b1, 1 - s, b2, s, b2, 1 - s] = j   # pair hopping
                um[b1, s, b2, 1 - s, b2, s, b1, 1 - s] = j   # spin flip
    return um


def _clebsch_gordan(j1, m1, j2, m2, jj, mm) -> float:
    """<j1 m1; j2 m2 | J M> for integer angular momenta (Racah's formula)."""
    if m1 + m2 != mm or abs(m1) > j1 or abs(m2) > j2 or abs(mm) > jj:
        return 0.0
    if jj > j1 + j2 or jj < abs(j1 - j2):
        return 0.0
    f = math.factorial
    pref = math.sqrt((2 * jj + 1) * f(jj + j1 - j2) * f(jj - j1 + j2) * f(j1 + j2 - jj)
                     / f(j1 + j2 + jj + 1))
    pref *= math.sqrt(f(jj + mm) * f(jj - mm) * f(j1 - m1) * f(j1 + m1)
                      * f(j2 - m2) * f(j2 + m2))
    s = 0.0
    for k in range(0, j1 + j2 - jj + 1):
        a = j1 + j2 - jj - k
        b = j1 - m1 - k
        c = j2 + m2 - k
        d = jj - j2 + m1 + k
        e = jj - j1 - m2 + k
        if min(a, b, c, d, e) < 0:
            continue
        s += (-1) ** k / (f(k) * f(a) * f(b) * f(c) * f(d) * f(e))
    return pref * s


def _gaunt_ak(k: int, l: int, m1: int, m2: int, m3: int, m4: int) -> float:
    """Angular factor a_k(m1 m2 m3 m4) of the Slater parametrisation."""
    a = sum(_clebsch_gordan(l, m3, k, q, l, m1) * _clebsch_gordan(l, m2, k, q, l, m4)
            for q in range(-k, k + 1))
    return a * _clebsch_gordan(l, 0, k, 0, l, 0) ** 2


def _cubic_rotation(nbands: int) -> np.ndarray:
    """Spherical -> cubic harmonics transformation with the orbital orders used by
    ``w2dyn/dmft/interaction.py:345-391`` (p: y,z,x; d: xy,yz,z2,xz,x2-y2)."""
    r = np.zeros((nbands, nbands), dtype=complex)
    s2 = 1.0 / math.sqrt(2.0)
    if nbands == 1:
        r[0, 0] = 1.0
    elif nbands == 3:
        r[0, 0], r[0, 2] = 1j * s2, 1j * s2
        r[1, 1] = 1.0
        r[2, 0], r[2, 2] = s2, -s2
    elif nbands == 5:
        r[0, 0], r[0, 4] = 1j * s2, -1j * s2
        r[1, 1], r[1, 3] = 1j * s2, 1j * s2
        r[2, 2] = 1.0
        r[3, 1], r[3, 3] = s2, -s2
        r[4, 0], r[4, 4] = s2, s2
    elif nbands == 7:
        r[0, 0], r[0, 6] = s2, -s2
        r[1, 1], r[1, 5] = s2, s2
        r[2, 2], r[2, 4] = s2, -s2
        r[3, 3] = 1.0
        r[4, 2], r[4, 4] = 1j * s2, 1j * s2
        r[5, 1], r[5, 5] = 1j * s2, -1j * s2
        r[6, 0], r[6, 6] = 1j * s2, 1j * s2
    else:
        raise NotImplementedError("Coulomb U only for 1, 3, 5 or 7 orbitals")
    return r


def u_coulomb(nbands: int, f0: float, f2: float = 0.0, f4: float = 0.0,
              f6: float = 0.0) -> np.ndarray:
    """Full rotationally invariant Coulomb U from Slater integrals in cubic
    harmonics (``w2dyn/dmft/interaction.py:275-410``)."""
    l = (nbands - 1) // 2
    ms = range(-l, l + 1)
    fk = {0: f0, 2: f2, 4: f4, 6: f6}
    usph = np.zeros((nbands,) * 4)
    for m1 in ms:
        for m2 in ms:
            for m3 in ms:
                for m4 in ms:
                    usph[m1 + l, m2 + l, m3 + l, m4 + l] = sum(
                        fk[k] * _gaunt_ak(k, l, m1, m2, m3, m4) for k in (0, 2, 4, 6) if k <= 2 * l)
    r = _cubic_rotation(nbands)
    ucub = np.einsum("ai,bj,ijkl,ck,dl->abcd", r.conj(), r.conj(), usph, r, r)
    um = np.zeros((nbands, 2) * 4)
    for s1 in (0, 1):
        for s2 in (0, 1):
            um[:, s1, :, s2, :, s1, :, s2] = ucub.real
    return um


def slater_from_uj_d(f0: float, j: float, f4_over_f2: float = 0.625):
    """d-shell Slater integrals from F0, Hund's J = (F2+F4)/14 and F4/F2."""
    f2 = 14.0 * j / (1.0 + f4_over_f2)
    return f0, f2, f4_over_f2 * f2


# --------------------------------------------------------------------------
# hybridisation
# --------------------------------------------------------------------------


def _fermi_kernel(tau: np.ndarray, eps: np.ndarray, beta: float) -> np.ndarray:
    """1/(exp(-tau e) + exp((beta - tau) e)) evaluated without overflow."""
    a = -tau[:, None] * eps[None, :]
    b = (beta - tau[:, None]) * eps[None, :]
    return np.exp(-np.logaddexp(a, b))


def semicircle_ftau(nbands: int, beta: float, half_bandwidth: float, nftau: int,
                    v2: float | None = None, nquad: int = 6000) -> np.ndarray:
    """Diagonal F(tau) of a semicircular bath of half-bandwidth D for every flavour.

    F(tau) = V^2 int de rho_D(e) / (exp(-tau e) + exp((beta - tau) e)) with
    rho_D the unit-weight semicircle and V^2 = D^2/4 unless given (the Bethe-lattice
    hybridisation).  Gauss-Chebyshev quadrature of the second kind.
    """
    d = float(half_bandwidth)
    if v2 is None:
        v2 = d * d / 4.0
    i = np.arange(1, nquad + 1)
    theta = i * np.pi / (nquad + 1)
    x = np.cos(theta)
    w = np.pi / (nquad + 1) * np.sin(theta) ** 2       # weights for sqrt(1-x^2)
    eps = d * x
    rhow = w * 2.0 / np.pi                              # unit-weight semicircle
    tau = np.linspace(0.0, beta, nftau)
    f = v2 * _fermi_kernel(tau, eps, beta) @ rhow
    ftau = np.zeros((nbands, 2, nbands, 2, nftau))
    for b in range(nbands):
        for s in range(2):
            ftau[b, s, b, s, :] = f
    return ftau


def discrete_bath_ftau(beta: float, eps: np.ndarray, vk: np.ndarray, nftau: int,
                       nbands: int = 1) -> np.ndarray:
    """F(tau) of discrete bath sites (``transform.py:145-146``), same for all flavours."""
    tau = np.linspace(0.0, beta, nftau)
    f = _fermi_kernel(tau, np.asarray(eps, float), beta) @ (np.asarray(vk, float) ** 2)
    ftau = np.zeros((nbands, 2, nbands, 2, nftau))
    for b in range(nbands):
        for s in range(2):
            ftau[b, s, b, s, :] = f
    return ftau


def solver_muimp(nbands: int, mu: float, crystalfield=None) -> np.ndarray:
    """One-particle term as passed to the solver: -(mu*1 - crystalfield)."""
    m = np.zeros((nbands, 2, nbands, 2))
    for b in range(nbands):
        for s in range(2):
            m[b, s, b, s] = mu
            if crystalfield is not None:
                m[b, s, b, s] -= crystalfield[b]
    return -m


def fock_hamiltonian(umatrix: np.ndarray, muimp: np.ndarray) -> np.ndarray:
    """Dense local Hamiltonian in the occupation basis of ``States.F90``: bit
    (b + N*s) of the state index holds the occupation of orbital b, spin s, and
    fermionic signs count occupied lower bits (``Operator.F90:115-200``)."""
    n = umatrix.shape[0]
    nf = 2 * n
    dim = 1 << nf

    def bit(b, s):
        return b + n * s

    def apply(op_bits, state):
        # op_bits: list of (bit, dagger) applied right to left
        sign = 1.0
        for bt, dag in reversed(op_bits):
            occ = (state >> bt) & 1
            if dag == occ:
                return None, 0.0
            if bin(state & ((1 << bt) - 1)).count("1") % 2:
                sign = -sign
            state ^= (1 << bt)
        return state, sign

    h = np.zeros((dim, dim))
    for b1 in range(n):
        for s1 in range(2):
            for b2 in range(n):
                for s2 in range(2):
                    val = muimp[b1, s1, b2, s2]
                    if val == 0:
                        continue
                    for st in range(dim):
                        t, sg = apply([(bit(b1, s1), 1), (bit(b2, s2), 0)], st)
                        if t is not None:
                            h[t, st] += val * sg
    idx = np.argwhere(umatrix != 0)
    for b1, s1, b2, s2, b3, s3, b4, s4 in idx:
        val = 0.5 * umatrix[b1, s1, b2, s2, b3, s3, b4, s4]
        ops = [(bit(b1, s1), 1), (bit(b2, s2), 1), (bit(b4, s4), 0), (bit(b3, s3), 0)]
        for st in range(dim):
            t, sg = apply(ops, st)
            if t is not None:
                h[t, st] += val * sg
    return h


def atomic_mu_window(umatrix: np.ndarray, muimp0: np.ndarray, filling: int) -> float:
    """Chemical potential in the middle of the window where the atomic ground
    state has ``filling`` electrons (``muimp0`` holds the mu-independent part)."""
    h = fock_hamiltonian(umatrix, muimp0)
    dim = h.shape[0]
    nel = np.array([bin(i).count("1") for i in range(dim)])
    emin = {}
    for nn in range(nel.max() + 1):
        sel = np.where(nel == nn)[0]
        emin[nn] = np.linalg.eigvalsh(h[np.ix_(sel, sel)])[0]
    lo = emin[filling] - emin[filling - 1]
    hi = emin[filling + 1] - emin[filling]
    return 0.5 * (lo + hi)


# --------------------------------------------------------------------------
# parameter string
# --------------------------------------------------------------------------

#: [QMC] defaults of ``w2dyn/auxiliaries/configspec:55-183`` that the sampling
#: path reads, plus the problem keys of ``cthyb_solver.py:226-237``.
QMC_DEFAULTS = {
    "Nwarmups": 10000, "Nmeas": 10000, "NCorr": 100, "NSeed": 43890, "NBin": 1,
    "MaxHisto": 500, "TaudiffMax": -1.0, "MeasDensityMatrix": 1,
    "PrintDensityMatrixBasis": 0, "MeasExpResDensityMatrix": 0, "MeasGiw": 0,
    "Gtau_mean_step": 0, "Gtau_mid_step": 0, "sign_step": 0, "NLookup_nfft": 10000,
    "EnableOuterTruncation": "NO", "truncation": 1, "EPSLANC": 1e-15,
    "EPSTRACEVEC": 1e-15, "EPSSANDWICH": 1e-15, "EPSEVEC": 1e-15, "EPSBLOCK": 1e-13,
    "PercentageGlobalMove": 0.005, "PercentageTauShiftMove": 0.005,
    "PercentageOuterMove": 0.0, "Ntau": 1000, "Nftau": 5000, "Niw": 2000,
    "NLegMax": 100, "NLegOrder": 100, "FourPnt": 0, "N4tau": 1, "N4leg": 2,
    "N4iwf": 1, "N4iwb": 0, "N2iwb": 100, "N3iwb": 100, "N3iwf": 100, "AccumGtau": 1,
    "MeasurementTiming": -1, "MeasG2iw": 0, "MeasG4iwPP": 0, "MeasGSigmaiw": 0,
    "MeasGtauDetRat": 0, "FixedPreallocatedVecs": 0, "Eigenbasis": 1,
    "statesampling": 0, "ReuseMCConfig": 0, "Update_Trace": 1, "offdiag": 0,
    "full_offdiag": 0, "flavourchange_moves": 0, "PercentageWormInsert": 0.0,
    "PercentageWormReplace": 0.0, "WormEta": 1.0, "WormSearchEta": 0,
    "WormPHConvention": 0, "ZPHConvention": 0, "G4ph": 1, "G4pp": 0,
    "nfft_mode_g4iw": 0, "Nbatchsize_nfft_worm": 5000000, "segment": 0,
    "MeasSusz": 0, "Percentage4OperatorMove": 0.0, "MeasSsquare": 0,
    "Nwarmups2Plus": -1, "real_ham_tolerance": 1e-10, "real_ftau_tolerance": 1e-10,
    "Nice_SelfEnergy": 0, "meas_dm_old": 0, "offd_strength": 0.0,
    "WormMeasGiw": 0, "WormMeasGtau": 0, "WormMeasGSigmaiw": 0, "WormMeasG4iw": 0,
    "WormMeasH4iw": 0, "WormMeasP2iwPH": 0, "WormMeasP2iwPP": 0, "WormMeasP2tauPH": 0,
    "WormMeasP2tauPP": 0, "WormMeasP3iwPH": 0, "WormMeasP3iwPP": 0, "WormMeasQQ": 0,
    "WormMeasQQtau": 0, "WormMeasNQQdag": 0, "WormMeasQQdd": 0, "WormMeasQQQQ": 0,
    "WormMeasUcaca": 0, "WormMeasUcacatau": 0, "WormMeasUccaa": 0,
    "WormMeasUccaatau": 0, "WormMeasQUDdag": 0,
}


@dataclass
class Problem:
    """One impurity problem plus the [QMC] settings of a solve."""
    nbands: int
    beta: float
    umatrix: np.ndarray
    ftau: np.ndarray
    muimp: np.ndarray
    quantum_numbers: str = "Nt Szt"
    interaction: str = "Density"
    params: dict = field(default_factory=dict)
    label: str = ""

    @property
    def nftau(self) -> int:
        return self.ftau.shape[-1]

    def qmc(self) -> dict:
        p = dict(QMC_DEFAULTS)
        p.update(self.params)
        p["Nftau"] = self.nftau
        return p

    def param_string(self) -> str:
        """Newline-separated ``key = value #`` string (``cthyb_solver.py:764-774``)."""
        q = self.qmc()
        s = "".join("%s = %s #\n" % (k, q[k]) for k in sorted(q))
        prob = {"Hamiltonian": self.interaction, "beta": self.beta, "Nd": self.nbands,
                "ParaMag": 0, "QuantumNumbers": self.quantum_numbers, "Phonon": 0,
                "Screening": 0, "Uw": 0}
        s += "".join("%s = %s #\n" % (k, prob[k]) for k in sorted(prob))
        if "NSymMove" not in q:
            s += "NSymMove = 0 #\n"
        return s

    def umatrix_flat(self) -> np.ndarray:
        """(2N)^4 view with flavour index 2*band+spin (``cthyb_solver.py:256``)."""
        n = 2 * self.nbands
        return np.ascontiguousarray(self.umatrix.reshape(n, n, n, n))


# --------------------------------------------------------------------------
# the benchmark configurations
# --------------------------------------------------------------------------


def anderson_1orb(beta=10.0, u=2.0, nftau=2001, **params) -> Problem:
    """1-orbital Anderson impurity at half filling, semicircular bath D=1."""
    p = {"Ntau": 1000, "NLegMax": 100}
    p.update(params)
    return Problem(1, beta, u_density(1, u, 0.0, 0.0), semicircle_ftau(1, beta, 1.0, nftau),
                   solver_muimp(1, u / 2.0), "Nt Szt", "Density", p, "1orb-anderson")


def kanamori_2orb(beta=50.0, u=4.0, j=1.0, nftau=2001, **params) -> Problem:
    """2-orbital Kanamori (U'=U-2J) at half filling, semicircular bath D=2."""
    v = u - 2 * j
    mu = (3 * u - 5 * j) / 2.0
    p = {"Ntau": 1000, "NLegMax": 100}
    p.update(params)
    return Problem(2, beta, u_kanamori(2, u, v, j), semicircle_ftau(2, beta, 2.0, nftau),
                   solver_muimp(2, mu), "Nt Szt", "Kanamori", p, "2orb-kanamori")


def kanamori_3orb(beta=100.0, u=3.5, j=0.6, mu=None, nftau=2001, seed=3, **params) -> Problem:
    """3-orbital t2g Kanamori with seeded crystal field in +-0.1, D=2."""
    v = u - 2 * j
    if mu is None:
        mu = 1.5 * (u - 3 * j)      # two electrons: middle of the atomic N=2 window
    cf = np.random.default_rng(seed).uniform(-0.1, 0.1, 3)
    p = {"Ntau": 1000, "NLegMax": 100}
    p.update(params)
    return Problem(3, beta, u_kanamori(3, u, v, j), semicircle_ftau(3, beta, 2.0, nftau),
                   solver_muimp(3, mu, cf), "Nt Szt", "Kanamori", p, "3orb-kanamori")


def coulomb_5orb(beta=40.0, f0=4.0, j=0.8, f4_over_f2=0.625, filling=6, mu=None,
                 nftau=2001, seed=5, **params) -> Problem:
    """5-orbital full Coulomb (Slater F0, F2, F4), seeded Gaussian crystal field."""
    f0, f2, f4 = slater_from_uj_d(f0, j, f4_over_f2)
    um = u_coulomb(5, f0, f2, f4)
    cf = np.random.default_rng(seed).normal(0.0, 0.2, 5)
    if mu is None:
        mu = atomic_mu_window(um, solver_muimp(5, 0.0, cf), filling)
    p = {"Ntau": 1000, "NLegMax": 100}
    p.update(params)
    return Problem(5, beta, um, semicircle_ftau(5, beta, 2.0, nftau),
                   solver_muimp(5, mu, cf), "Nt Szt", "Coulomb", p, "5orb-coulomb")


# --------------------------------------------------------------------------
# parameter derivatives of the solver inputs
# --------------------------------------------------------------------------


@dataclass
class InputDerivatives:
    """Derivatives of the ``init_solver`` inputs with respect to model
    parameters theta: ``du[t]`` of the U matrix (N,2,N,2,N,2,N,2), ``dmuimp[t]``
    of the one-particle term (N,2,N,2), ``dftau[t]`` of F(tau) (N,2,N,2,Nftau)."""
    names: tuple
    du: np.ndarray
    dmuimp: np.ndarray
    dftau: np.ndarray

    @property
    def ntheta(self) -> int:
        return len(self.names)


#: the parameters of the differentiable 3-orbital Kanamori model
KANAMORI_THETA = ("U", "J", "eps_d", "V")


def kanamori_3orb_theta(theta=(3.5, 0.6, 0.0, 1.0), beta=100.0, mu=2.55, half_bandwidth=2.0,
                        nftau=2001, seed=3, **params):
    """The 3-orbital Kanamori problem as a function of theta = (U, J, eps_d, V),
    and the derivatives of its solver inputs.

    * U, J: Kanamori interaction with U' = U - 2J (``interaction.py:221-244``);
      the U matrix is linear in (U, J).
    * eps_d: a common shift of the three orbital levels on top of the seeded
      crystal field; the one-particle term gains eps_d on the diagonal.
    * V: bath coupling, F(tau) = V^2 int de rho_D(e) / (exp(-tau e) +
      exp((beta - tau) e)), so dF/dV = 2 F / V (V = D/2 is the Bethe value).
    * mu is held fixed (2.55 = 1.5 (U - 3J) at U = 3.5, J = 0.6).
    """
    u, j, eps, v = (float(x) for x in theta)
    cf = np.random.default_rng(seed).uniform(-0.1, 0.1, 3)
    p = {"Ntau": 1000, "NLegMax": 100}
    p.update(params)
    ftau1 = semicircle_ftau(3, beta, half_bandwidth, nftau, v2=1.0)
    prob = Problem(3, beta, u_kanamori(3, u, u - 2 * j, j), v * v * ftau1,
                   solver_muimp(3, mu, cf + eps), "Nt Szt", "Kanamori", p, "3orb-kanamori-theta")
    nb = 3
    du = np.stack([u_kanamori(nb, 1.0, 1.0, 0.0), u_kanamori(nb, 0.0, -2.0, 1.0),
                   np.zeros((nb, 2) * 4), np.zeros((nb, 2) * 4)])
    dmu = np.zeros((4, nb, 2, nb, 2))
    for b in range(nb):
        for s in range(2):
            dmu[2, b, s, b, s] = 1.0
    dft = np.zeros((4,) + ftau1.shape)
    dft[3] = 2.0 * v * ftau1
    return prob, InputDerivatives(KANAMORI_THETA, du, dmu, dft)


#: the parameters of the differentiable 1-orbital Anderson model
ANDERSON_THETA = ("U", "eps_d", "V")


def anderson_discrete_theta(theta=(2.0, -0.6, 1.0), beta=5.0, eps_k=(-1.0, 0.2, 1.1),
                            v_k=(0.5, 0.35, 0.45), nftau=4001, **params):
    """1-orbital Anderson impurity with discrete bath sites as a function of
    theta = (U, eps_d, V): H_loc = eps_d (n_up + n_dn) + U n_up n_dn, bath
    couplings V v_k (F ~ V^2, dF/dV = 2 F / V); exactly solvable by
    diagonalising impurity + bath."""
    u, eps, v = (float(x) for x in theta)
    p = {"Ntau": 1000, "NLegMax": 100}
    p.update(params)
    f1 = discrete_bath_ftau(beta, np.asarray(eps_k, float), np.asarray(v_k, float), nftau)
    prob = Problem(1, beta, u_density(1, u, 0.0, 0.0), v * v * f1, solver_muimp(1, -eps), "Nt Szt",
                   "Density", p, "1orb-anderson-discrete-theta")
    du = np.stack([u_density(1, 1.0, 0.0, 0.0), np.zeros((1, 2) * 4), np.zeros((1, 2) * 4)])
    dmu = np.zeros((3, 1, 2, 1, 2))
    dmu[1, 0, 0, 0, 0] = dmu[1, 0, 1, 0, 1] = 1.0
    dft = np.zeros((3,) + f1.shape)
    dft[2] = 2.0 * v * f1
    return prob, InputDerivatives(ANDERSON_THETA, du, dmu, dft)
