"""ctypes front end of the C oracle (``ctqmc_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, never by the product.  ``OracleChain`` mirrors
one Markov chain of the reference module ``MCTQMC`` (``CTQMC.F90``): creation
does what ``init_rnd`` + ``init_solver`` do, ``moves`` what the ``ctqmc_warmup``
loop does, ``measure`` what ``ctqmc_measure`` does, and ``results`` applies the
normalisation of ``CTQMC.F90:8663-8771``.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from model import OracleModel  # noqa: F401  (re-exported for tests)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libctqmc_oracle.so")

MOVE_DT = np.dtype([("kind", "<i4"), ("outcome", "<i4"), ("noper", "<i4"), ("outer_sst", "<i4"),
                    ("window", "<i4"), ("trace_log", "<f8"), ("trace_sign", "<f8"),
                    ("det_log", "<f8"), ("det_sign", "<f8"), ("cfgsign", "<f8")], align=True)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        P = C.c_void_p
        _lib.oracle_create.restype = P
        _lib.oracle_create.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, C.c_double, P, P, P, P,
                                       P, P, P, C.c_int, P, C.c_double, P, P, C.c_uint]
        _lib.oracle_destroy.argtypes = [P]
        _lib.oracle_set_taudiffmax.argtypes = [P, C.c_double]
        _lib.oracle_moves.argtypes = [P, C.c_long, P]
        _lib.oracle_reset_counters.argtypes = [P]
        _lib.oracle_measure.argtypes = [P, C.c_long, C.c_long]
        _lib.oracle_accumulators.argtypes = [P, P, P, P, P, P, P]
        _lib.oracle_get_config.argtypes = [P, C.c_int, P, P, P, P, P, P]
        _lib.oracle_get_config.restype = C.c_int
        _lib.oracle_get_minv.argtypes = [P, C.c_int, P]
        _lib.oracle_get_minv.restype = C.c_int
        _lib.oracle_set_config.argtypes = [P, C.c_int, P, P, P, P, C.c_int, P]
        _lib.oracle_rng_stream.argtypes = [C.c_uint, C.c_long, P, P, P, P]
        _lib.oracle_giw.argtypes = [P, P]
        _lib.oracle_giw.restype = C.c_int
        _lib.oracle_gtau_blocks.argtypes = [P, P]
        _lib.oracle_gtau_blocks.restype = C.c_int
        _lib.oracle_autocorr.argtypes = [C.c_long, C.c_int, C.c_long, P, P, P]
        _lib.oracle_begin_warmup.argtypes = [P]
        _lib.oracle_calibrate.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        _lib.oracle_calibrate.restype = C.c_long
        _lib.oracle_pair_stats.argtypes = [P, P, P, P]
        _lib.oracle_pair_stats.restype = C.c_int
        _lib.oracle_outer_histos.argtypes = [P, P, P, P, P, P]
        _lib.oracle_steps.argtypes = [P, P, P, P]
        _lib.oracle_steps.restype = C.c_int
        _lib.oracle_set_symmoves.argtypes = [P, C.c_int, P]
        _lib.oracle_set_derivatives.argtypes = [P, C.c_int, P, P, C.c_int]
        _lib.oracle_set_derivatives.restype = C.c_int
        _lib.oracle_dlogw.argtypes = [P, P]
        _lib.oracle_score_sums.argtypes = [P, P, P, P, P]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def autocorr(data: np.ndarray, nlev: int):
    """The binning accumulator (AccumGtau = 2) on a sequence data[n][ncomp]:
    (mean, blocks[nlev][ncomp])."""
    data = np.ascontiguousarray(data, dtype=float)
    n, nc = data.shape
    mean = np.zeros(nc)
    blocks = np.zeros((nlev, nc))
    lib().oracle_autocorr(nc, int(nlev), n, _p(data), _p(mean), _p(blocks))
    return mean, blocks


def rng_stream(seed: int, kinds, lo=None, hi=None) -> np.ndarray:
    """Draws of the reference RNG: kind 0 = grnd(), 1 = randint(lo, hi), 2 = raw."""
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    n = len(kinds)
    lo = np.zeros(n, np.int32) if lo is None else np.ascontiguousarray(lo, np.int32)
    hi = np.zeros(n, np.int32) if hi is None else np.ascontiguousarray(hi, np.int32)
    out = np.zeros(n)
    lib().oracle_rng_stream(seed, n, _p(kinds), _p(lo), _p(hi), _p(out))
    return out


class OracleChain:
    """One reference-faithful Markov chain on the CPU."""

    def __init__(self, umatrix, ftau, muimp, beta, qmc: dict, seed: int,
                 quantum_numbers: str | None = None):
        if quantum_numbers is None:
            quantum_numbers = str(qmc.get("QuantumNumbers", "Nt Szt"))
        self.model = OracleModel(np.asarray(umatrix), np.asarray(muimp), quantum_numbers,
                                 float(qmc.get("EPSBLOCK", 1e-13)))
        m = self.model
        fl = m.flat()
        nb = m.nbands
        self.nb, self.beta = nb, float(beta)
        self.qmc = dict(qmc)
        # keep every array alive for the lifetime of the chain
        ftau = np.asarray(ftau)
        fdiag = np.zeros((2 * nb, ftau.shape[-1]))
        for b in range(nb):
            for s in range(2):
                fdiag[b + nb * s] = ftau[b, s, b, s]
        self._keep = dict(fl)
        self._keep["ftau"] = np.ascontiguousarray(fdiag)
        self._keep["dparams"] = np.array([float(qmc["PercentageTauShiftMove"]),
                                          float(qmc["PercentageGlobalMove"]),
                                          float(qmc["PercentageOuterMove"]),
                                          float(qmc["Percentage4OperatorMove"]),
                                          float(qmc["TaudiffMax"])])
        self._keep["iparams"] = np.array([int(qmc["NCorr"]), int(qmc["Ntau"]), int(qmc["NLegMax"]),
                                          int(qmc["MaxHisto"]),
                                          int("Szt" in quantum_numbers),
                                          int(qmc["MeasDensityMatrix"]),
                                          int(qmc.get("AccumGtau", 1)),
                                          int(qmc["Nmeas"]),
                                          int(qmc.get("Gtau_mean_step", 0)),
                                          int(qmc.get("Gtau_mid_step", 0)),
                                          int(qmc.get("sign_step", 0)),
                                          int(qmc.get("Niw", 2000)) if int(qmc.get("MeasGiw", 0)) else 0],
                                         dtype=np.int64)
        for key in ("opoff", "evoff"):
            self._keep[key] = np.ascontiguousarray(self._keep[key], dtype=np.int64)
        k = self._keep
        self.ntau, self.nleg = int(qmc["Ntau"]), int(qmc["NLegMax"])
        self.maxhisto = int(qmc["MaxHisto"])
        self.h = lib().oracle_create(nb, fl["nsst"], fl["nmax"], _p(k["sizes"]), _p(k["offsets"]),
                                     _p(k["evals"]), float(fl["egs"]), _p(k["connect"]),
                                     _p(k["opoff"]), _p(k["ops"]), _p(k["evec"]), _p(k["evoff"]),
                                     _p(k["first_state"]), _p(k["state_sst"]),
                                     int(k["ftau"].shape[1]), _p(k["ftau"]), float(beta),
                                     _p(k["dparams"]), _p(k["iparams"]), int(seed) & 0xffffffff)

        nsym = int(float(qmc.get("NSymMove", 0)))
        if nsym > 0:
            perms = np.array([[int(x) for x in str(qmc["SymMove%02d" % (i + 1)]).replace(",", " ").split()]
                              for i in range(nsym)], dtype=np.int32)
            self._keep["symmoves"] = np.ascontiguousarray(perms)
            lib().oracle_set_symmoves(self.h, nsym, _p(self._keep["symmoves"]))

    @classmethod
    def from_problem(cls, prob, seed: int):
        return cls(prob.umatrix, prob.ftau, prob.muimp, prob.beta, prob.qmc(), seed,
                   prob.quantum_numbers)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    # -- sampling ------------------------------------------------------------
    def set_taudiffmax(self, tdm: float):
        lib().oracle_set_taudiffmax(self.h, float(tdm))

    def moves(self, n: int, record: bool = True):
        rec = np.zeros(n, dtype=MOVE_DT) if record else None
        lib().oracle_moves(self.h, int(n), _p(rec) if record else None)
        return rec

    def warmup(self, n: int):
        lib().oracle_begin_warmup(self.h)
        lib().oracle_moves(self.h, int(n), None)

    def init_counters(self):
        lib().oracle_reset_counters(self.h)

    def calibrate(self, list_phase: bool, phase1pairnum=1000, phase2taugrid=100, phase2pairnum=1000,
                  nprogress=100000) -> int:
        """ctqmc_calibrate; returns the moves it took."""
        return lib().oracle_calibrate(self.h, 1 if list_phase else 0, int(phase1pairnum), int(phase2taugrid),
                                      int(phase2pairnum), int(nprogress))

    def pair_stats(self) -> dict:
        nb = C.c_int()
        tau = np.zeros(1 << 16)
        ap = np.zeros(2 * self.nb * 8192)
        n = lib().oracle_pair_stats(self.h, _p(tau), _p(ap), C.byref(nb))
        return dict(accpairtau=tau[:n].copy(),
                    accpair=ap[:2 * self.nb * nb.value].reshape(2 * self.nb, nb.value).copy())

    def outer_histos(self) -> dict:
        nst, nss = 4 ** self.nb, self.model.nsst
        out = dict(outer_state=np.zeros(nst), sign_sst=np.zeros(nss), sign_state=np.zeros(nst),
                   contrib_sst=np.zeros(nss), contrib_state=np.zeros(nst))
        lib().oracle_outer_histos(self.h, *[_p(out[k]) for k in ("outer_state", "sign_sst", "sign_state",
                                                                  "contrib_sst", "contrib_state")])
        return out

    def steps(self) -> dict:
        n = lib().oracle_steps(self.h, None, None, None)
        nf = 2 * self.nb
        m, d, sg = np.zeros((max(n, 1), nf)), np.zeros((max(n, 1), nf)), np.zeros(max(n, 1))
        lib().oracle_steps(self.h, _p(m), _p(d), _p(sg))
        return dict(gtau_mean_step=m[:n], gtau_mid_step=d[:n], sign_step=sg[:n])

    def measure(self, nmeas: int, ncorr: int):
        lib().oracle_reset_counters(self.h)
        lib().oracle_measure(self.h, int(nmeas), int(ncorr))

    # -- inspection ------------------------------------------------------------
    def config(self):
        cap = 1 << 16
        tau = np.zeros(cap)
        orb, spin, ca = (np.zeros(cap, np.int32) for _ in range(3))
        outer = np.zeros(1, np.int32)
        w = np.zeros(6)
        n = lib().oracle_get_config(self.h, cap, _p(tau), _p(orb), _p(spin), _p(ca), _p(outer), _p(w))
        return dict(n=n, taus=tau[:n].copy(), orbs=orb[:n].copy(), spins=spin[:n].copy(),
                    cas=ca[:n].copy(), outer_sst=int(outer[0]), trace_log=w[0], trace_sign=w[1],
                    det_log=w[2], det_sign=w[3], cfgsign=w[4], window=int(w[5]))

    def minv(self, f: int) -> np.ndarray:
        k = lib().oracle_get_minv(self.h, f, None)
        out = np.zeros(max(k * k, 1))
        lib().oracle_get_minv(self.h, f, _p(out))
        return out[:k * k].reshape(k, k, order="F")

    def set_config(self, taus, orbs, spins, cas, outer_sst):
        """0-based orbitals/spins, ca 1 = creator, 2 = annihilator."""
        taus = np.ascontiguousarray(taus, float)
        orbs, spins, cas = (np.ascontiguousarray(x, np.int32) for x in (orbs, spins, cas))
        w = np.zeros(6)
        lib().oracle_set_config(self.h, len(taus), _p(taus), _p(orbs), _p(spins), _p(cas),
                                int(outer_sst), _p(w))
        return dict(trace_log=w[0], trace_sign=w[1], det_log=w[2], det_sign=w[3], cfgsign=w[4])

    # -- parameter derivatives (no reference counterpart) ----------------------
    def set_derivatives(self, deriv, score: bool = True):
        """Derivative tables of the inputs (an InputDerivatives: du, dmuimp,
        dftau per theta); with ``score`` every measurement adds the
        score-function sums."""
        nt = len(deriv.du)
        blk = np.concatenate([self.model.derivative_blocks(deriv.du[t], deriv.dmuimp[t])
                              for t in range(nt)])
        nb = self.nb
        dft = np.zeros((nt, 2 * nb, deriv.dftau.shape[-1]))
        for t in range(nt):
            for b in range(nb):
                for sp in range(2):
                    dft[t, b + nb * sp] = deriv.dftau[t, b, sp, b, sp]
        self._keep["dblk"] = np.ascontiguousarray(blk)
        self._keep["dftau_d"] = np.ascontiguousarray(dft)
        self.ntheta = nt
        st = lib().oracle_set_derivatives(self.h, nt, _p(self._keep["dblk"]), _p(self._keep["dftau_d"]),
                                          int(bool(score)))
        assert st == 0

    def dlogw(self) -> np.ndarray:
        """[theta][0: d ln|det|, 1: d ln|Tr|] of the current configuration."""
        out = np.zeros(2 * self.ntheta)
        lib().oracle_dlogw(self.h, _p(out))
        return out.reshape(self.ntheta, 2)

    def score_sums(self) -> dict:
        """Sum sign s, Sum s G-sample [theta][f][tau], Sum s rho-sample and
        Sum d(rho-sample)/dtheta [theta][state][state] (eigenbasis, dense)."""
        nt, nf, nst = self.ntheta, 2 * self.nb, 4 ** self.nb
        ss = np.zeros(nt)
        g = np.zeros(nt * nf * self.ntau)
        dm = np.zeros(nt * nst * nst)
        dmx = np.zeros(nt * nst * nst)
        lib().oracle_score_sums(self.h, _p(ss), _p(g), _p(dm), _p(dmx))
        return dict(ssum=ss, gtau=g.reshape(nt, nf, self.ntau),
                    dm=dm.reshape(nt, nst, nst).transpose(0, 2, 1).copy(),
                    dmx=dmx.reshape(nt, nst, nst).transpose(0, 2, 1).copy())

    def gtau_blocks(self) -> np.ndarray:
        """AccumGtau = 2: variances of the level-l bin means, [nlev][flavour][ntau]."""
        nlev = lib().oracle_gtau_blocks(self.h, None)
        out = np.zeros(nlev * 2 * self.nb * self.ntau)
        lib().oracle_gtau_blocks(self.h, _p(out))
        return out.reshape(nlev, 2 * self.nb, self.ntau)

    def giw_raw(self) -> np.ndarray:
        """MeasGiw: raw sums of cfgsign M(c,a) exp(i w_n (tau_a - tau_c)) over the
        measurements, [flavour][n] for n = 0..Niw-1 (complex)."""
        niw = lib().oracle_giw(self.h, None)
        out = np.zeros(2 * self.nb * max(niw, 1) * 2)
        lib().oracle_giw(self.h, _p(out))
        out = out[:2 * self.nb * niw * 2].reshape(2 * self.nb, niw, 2)
        return out[..., 0] + 1j * out[..., 1]

    def raw(self) -> dict:
        nf = 2 * self.nb
        nst = 4 ** self.nb
        g = np.zeros(nf * self.ntau)
        gl = np.zeros(nf * self.nleg)
        dm = np.zeros(nst * nst)
        hi = np.zeros(nf * (self.maxhisto + 1))
        oh = np.zeros(self.model.nsst)
        sc = np.zeros(10)
        lib().oracle_accumulators(self.h, _p(g), _p(gl), _p(dm), _p(hi), _p(oh), _p(sc))
        return dict(gtau=g.reshape(nf, self.ntau), gleg=gl.reshape(nf, self.nleg),
                    dm=dm.reshape(nst, nst, order="F"), histo=hi.reshape(nf, self.maxhisto + 1),
                    outer_hist=oh, scalars=sc)

    def results(self) -> dict:
        """Normalised observables as in ctqmc_measure's post-processing."""
        r = self.raw()
        nb, beta = self.nb, self.beta
        sc = r["scalars"]
        cnt = sc[1]
        mean_sign = sc[0] / cnt
        zs = -1.0 / (cnt * mean_sign)

        def fl2(a):   # [f, ...] -> [b, s, ...]
            return a.reshape((2, nb) + a.shape[1:]).swapaxes(0, 1)

        gtau = fl2(r["gtau"]) / cnt
        gtau = -gtau / (-mean_sign) * self.ntau / beta ** 2
        gleg = fl2(r["gleg"]).copy()
        for l in range(self.nleg):
            gleg[..., l] *= np.sqrt(2.0 * l + 1.0) / beta
        gleg = -gleg * zs
        dm_eb = r["dm"] * zs
        m = self.model
        # occupation basis (Transform_DensityMatrix_EB_to_OCCB, CTQMC.F90:3163-3179)
        dm_occ = dm_eb.copy()
        for s in range(m.nsst):
            o, n = m.offsets[s], m.sizes[s]
            v = m.evecs[s]
            dm_occ[o:o + n, o:o + n] = v @ dm_eb[o:o + n, o:o + n] @ v.T
        # occ(b1,s1,b2,s2) = <n_{b2 s2} n_{b1 s1}> from the diagonal in the
        # occupation basis (MeasDoubleOcc/SingleOcc_from_Densitymatrix)
        states = np.concatenate(m.states)
        diag = np.diag(dm_occ)
        occ = np.zeros((nb, 2, nb, 2))
        for b1 in range(nb):
            for s1 in range(2):
                n1 = (states >> (b1 + nb * s1)) & 1
                for b2 in range(nb):
                    for s2 in range(2):
                        n2 = (states >> (b2 + nb * s2)) & 1
                        occ[b1, s1, b2, s2] = np.sum(diag * n1 * n2)
        acc = np.array([sc[2] / sc[3] if sc[3] else np.nan, sc[4] / sc[5] if sc[5] else np.nan,
                        sc[6] / sc[7] if sc[7] else np.nan, sc[8] / sc[9] if sc[9] else np.nan])
        out = dict(mean_sign=mean_sign, gtau=gtau, gleg=gleg, occ=occ, densitymatrix=dm_occ,
                   histo=fl2(r["histo"]) / cnt, acc=acc, outer_sst_histo=r["outer_hist"] / cnt)
        gw = self.giw_raw()
        if gw.shape[1]:
            # Giw * zScaling / beta (CTQMC.F90:8816) on the sums of the lookup
            # path's -cfgsign M terms; n = -Niw..Niw-1, n < 0 the conjugates
            pos = fl2(gw) / (cnt * mean_sign * beta)
            out["giw"] = np.concatenate([np.conj(pos[..., ::-1]), pos], axis=-1)
        return out
