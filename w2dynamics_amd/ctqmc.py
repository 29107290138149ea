"""Python host of the MI355X CT-HYB sampling core.

``CtqmcGpu`` drives one handle of the C-ABI (``include/w2dyn_ctqmc.h``), i.e. a
batch of independent Markov chains on one GPU.  Its methods mirror the entry
points of the reference solver module ``MCTQMC`` that the Python driver calls
(``w2dyn/dmft/cthyb_solver.py:240-355``):

=======================  ==================================================
reference (CTQMC.F90)    here
=======================  ==================================================
init_paras + init_rnd    ``CtqmcGpu(params, ..., seed=NSeed)``
+ init_solver            (chain c is seeded NSeed + c)
set_taudiffmax           ``set_taudiffmax``
ctqmc_warmup             ``warmup``
ctqmc_measure(1, 1)      ``measure``
get/set_mc_config        ``get_config`` / ``set_config``
module arrays            ``results()`` (per chain) and ``aggregate()``
dest_ctqmc               ``close``
=======================  ==================================================

The normalisation applied in ``results`` is the one of ``CTQMC.F90:8663-8771``;
``aggregate`` combines chains like ``DistributedSample`` combines MPI ranks
(``w2dyn/auxiliaries/statistics.py:130-209``).
"""

from __future__ import annotations

import ctypes as C
import os
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBNAME = os.path.join(HERE, "libw2dyn_ctqmc.so")

W2D_ERRORS = {1: "invalid argument or unsupported parameter", 2: "HIP runtime error",
              3: "capacity exceeded (expansion order or window length)",
              4: "DetRat too small in removal", 5: "zero trace initial state",
              6: "call out of order", 7: "some chains stopped on an error"}
W2D_ERR_PARTIAL = 7


class CtqmcError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__("%s failed: %s (code %d)" % (where, W2D_ERRORS.get(code, "?"), code))
        self.code = code


MOVE_DT = np.dtype([("kind", "<i4"), ("outcome", "<i4"), ("noper", "<i4"), ("outer_sst", "<i4"),
                    ("window", "<i4"), ("pad", "<i4"), ("trace_log", "<f8"), ("trace_sign", "<f8"),
                    ("det_log", "<f8"), ("det_sign", "<f8"), ("cfgsign", "<f8")])

_lib = None


def load_library(path: str | None = None):
    """Load the HIP engine; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIBNAME
    if not os.path.exists(path):
        raise ImportError("libw2dyn_ctqmc.so not built: run __graft_entry__.build() "
                          "(expected at %s)" % path)
    lib = C.CDLL(path)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    lib.w2d_ctqmc_create.argtypes = [C.c_char_p, I, I, P, P, P, I, C.c_uint32, I, I, C.POINTER(P)]
    lib.w2d_ctqmc_set_taudiffmax.argtypes = [P, D]
    lib.w2d_ctqmc_warmup.argtypes = [P, C.c_int64]
    lib.w2d_ctqmc_measure.argtypes = [P, C.c_int64]
    lib.w2d_ctqmc_dims.argtypes = [P] + [C.POINTER(I)] * 7
    lib.w2d_ctqmc_get_accumulators.argtypes = [P] * 7
    lib.w2d_ctqmc_get_model.argtypes = [P] * 7
    lib.w2d_ctqmc_moves_log.argtypes = [P, I, C.c_int64, P]
    lib.w2d_ctqmc_get_config.argtypes = [P, I, I, C.POINTER(I), P, P, P, P, C.POINTER(I), P]
    lib.w2d_ctqmc_set_config.argtypes = [P, I, I, P, P, P, P, I, P]
    lib.w2d_ctqmc_get_minv.argtypes = [P, I, I, C.POINTER(I), P]
    lib.w2d_ctqmc_last_kernel_ms.argtypes = [P, C.POINTER(D), C.POINTER(C.c_int64)]
    lib.w2d_ctqmc_destroy.argtypes = [P]
    lib.w2d_ctqmc_get_group_sums.argtypes = [P, I, P, P, P, P, P, P, P]
    lib.w2d_ctqmc_chain_status.argtypes = [P, P, P, C.POINTER(I)]
    lib.w2d_ctqmc_set_abort.argtypes = [P, I]
    lib.w2d_ctqmc_aborted.argtypes = [P, C.POINTER(I)]
    lib.w2d_ctqmc_trap_signals.argtypes = []
    lib.w2d_ctqmc_get_gtau_blocks.argtypes = [P, I, C.POINTER(I), P]
    lib.w2d_ctqmc_get_giw.argtypes = [P, I, C.POINTER(I), P]
    lib.w2d_ctqmc_get_orders.argtypes = [P, P, P]
    lib.w2d_ctqmc_op_table_bytes.argtypes = [P, C.POINTER(C.c_int64)]
    lib.w2d_ctqmc_init_counters.argtypes = [P]
    lib.w2d_ctqmc_calibrate.argtypes = [P, I, I, I, I, I]
    lib.w2d_ctqmc_get_pair_stats.argtypes = [P, I, P, C.POINTER(I), P, C.POINTER(I)]
    lib.w2d_ctqmc_get_group_histograms.argtypes = [P, I, P, P, P, P, P, P]
    lib.w2d_ctqmc_get_steps.argtypes = [P, I, C.POINTER(I), P, P, P]
    lib.w2d_ctqmc_get_group_sums_noalloc.argtypes = [P, I, P, P, P, P, P, P, P]
    lib.w2d_ctqmc_set_launch.argtypes = [P, I, I, C.c_int64]
    lib.w2d_ctqmc_set_engine.argtypes = [P, I]
    lib.w2d_ctqmc_get_engine.argtypes = [P, C.POINTER(I), C.POINTER(I)]
    lib.w2d_ctqmc_autotune.argtypes = [P, C.c_int64, C.POINTER(I), P]
    lib.w2d_ctqmc_set_derivatives.argtypes = [P, I, P, P, P, I]
    lib.w2d_ctqmc_dlogw.argtypes = [P, I, P]
    lib.w2d_ctqmc_get_score_sums.argtypes = [P, I, P, P, P, P]
    lib.w2d_ctqmc_version.restype = C.c_char_p
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(code: int, where: str):
    if code != 0:
        raise CtqmcError(code, where)


class CtqmcGpu:
    """A batch of ``nchains`` CT-HYB Markov chains on one GPU."""

    def __init__(self, params: str, umatrix, ftau, muimp, nchains: int = 1, seed: int = 43890,
                 capacity: int = 256, device: int = 0):
        lib = load_library()
        umatrix = np.asarray(umatrix, dtype=np.float64)
        nb = int(np.asarray(muimp).shape[0])
        n2 = 2 * nb
        u = np.asfortranarray(umatrix.reshape(n2, n2, n2, n2))
        f = np.asfortranarray(np.asarray(ftau, dtype=np.float64))
        m = np.asfortranarray(np.asarray(muimp, dtype=np.float64))
        self._inputs = (u, f, m)
        h = C.c_void_p()
        _check(lib.w2d_ctqmc_create(params.encode(), nb, f.shape[-1], _p(u), _p(f), _p(m),
                                    int(nchains), int(seed) & 0xffffffff, int(capacity), int(device),
                                    C.byref(h)), "w2d_ctqmc_create")
        self.h = h
        self.lib = lib
        dims = [C.c_int() for _ in range(7)]
        _check(lib.w2d_ctqmc_dims(h, *[C.byref(x) for x in dims]), "w2d_ctqmc_dims")
        (self.nbands, self.nchains, self.ntau, self.nleg, self.nsst, self.nstates,
         self.maxhisto) = [x.value for x in dims]
        self.beta = _param_value(params, "beta")
        self._model = None

    @classmethod
    def from_problem(cls, prob, nchains: int = 1, seed: int = 43890, capacity: int = 256,
                     device: int = 0):
        return cls(prob.param_string(), prob.umatrix, prob.ftau, prob.muimp, nchains, seed,
                   capacity, device)

    def close(self):
        if getattr(self, "h", None):
            self.lib.w2d_ctqmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- sampling ---------------------------------------------------------------
    def set_taudiffmax(self, x: float):
        _check(self.lib.w2d_ctqmc_set_taudiffmax(self.h, float(x)), "w2d_ctqmc_set_taudiffmax")

    def warmup(self, n: int):
        self._check_batch(self.lib.w2d_ctqmc_warmup(self.h, int(n)), "w2d_ctqmc_warmup")

    def measure(self, nmeas: int):
        self._check_batch(self.lib.w2d_ctqmc_measure(self.h, int(nmeas)), "w2d_ctqmc_measure")

    def _check_batch(self, code: int, where: str):
        """A batch where some chains stopped on an error carries on with the
        others (their results are left out, ``chain_status`` names them)."""
        if code == W2D_ERR_PARTIAL:
            st = self.chain_status()["status"]
            warnings.warn("%s: %d of %d chains stopped on an error (codes %s); their results are "
                          "left out" % (where, int((st != 0).sum()), self.nchains,
                                        sorted(set(st[st != 0].tolist()))), RuntimeWarning)
            return
        _check(code, where)

    def chain_status(self) -> dict:
        """Per-chain status (0 = ok, else the error that stopped it), the largest
        order each flavour reached, and the current pairs-per-flavour capacity."""
        st = np.zeros(self.nchains, np.int32)
        km = np.zeros((self.nchains, 2 * self.nbands), np.int32)
        cap = C.c_int()
        _check(self.lib.w2d_ctqmc_chain_status(self.h, _p(st), _p(km), C.byref(cap)), "chain_status")
        return dict(status=st, kfmax=km, capacity=cap.value)

    def init_counters(self):
        """init_counters: move counters, AccPair and the calibration list index."""
        _check(self.lib.w2d_ctqmc_init_counters(self.h), "init_counters")

    def calibrate(self, list_phase: bool, phase1pairnum: int = 1000, phase2taugrid: int = 100,
                  phase2pairnum: int = 1000, nprogress: int = 100000):
        """ctqmc_calibrate for every chain (each stops on its own condition)."""
        self._check_batch(self.lib.w2d_ctqmc_calibrate(self.h, 1 if list_phase else 0, int(phase1pairnum),
                                                       int(phase2taugrid), int(phase2pairnum),
                                                       int(nprogress)), "calibrate")

    def pair_stats(self, chain: int = 0) -> dict:
        """AccPairTau (distances of calibration phase 1) and AccPair [flavour][bin]."""
        n, nb = C.c_int(), C.c_int()
        _check(self.lib.w2d_ctqmc_get_pair_stats(self.h, chain, None, C.byref(n), None, C.byref(nb)),
               "pair_stats")
        tau = np.zeros(max(n.value, 1))
        ap = np.zeros((2 * self.nbands, max(nb.value, 1)))
        _check(self.lib.w2d_ctqmc_get_pair_stats(self.h, chain, _p(tau), C.byref(n), _p(ap), C.byref(nb)),
               "pair_stats")
        return dict(accpairtau=tau[:n.value], accpair=ap[:, :nb.value])

    def group_histograms(self, ngroups: int) -> dict:
        """Outer-state, sign and trace-contribution histograms and AccPair summed
        over groups of chains (raw)."""
        ng, nst, nss = int(ngroups), self.nstates, self.nsst
        cap = max(self.ntau, 1024)
        out = dict(outer_state=np.zeros((ng, nst)), sign_sst=np.zeros((ng, nss)),
                   sign_state=np.zeros((ng, nst)), contrib_sst=np.zeros((ng, nss)),
                   contrib_state=np.zeros((ng, nst)), accpair=np.zeros((ng, 2 * self.nbands, cap)))
        _check(self.lib.w2d_ctqmc_get_group_histograms(
            self.h, ng, *[_p(out[k]) for k in ("outer_state", "sign_sst", "sign_state", "contrib_sst",
                                              "contrib_state", "accpair")]), "group_histograms")
        return out

    def steps(self, chain: int = 0) -> dict:
        """Gtau_mean_step, Gtau_mid_step [Nmeas][flavour] and sign_step [Nmeas] (raw)."""
        n = C.c_int()
        _check(self.lib.w2d_ctqmc_get_steps(self.h, chain, C.byref(n), None, None, None), "steps")
        nf = 2 * self.nbands
        m, d, sg = np.zeros((max(n.value, 1), nf)), np.zeros((max(n.value, 1), nf)), np.zeros(max(n.value, 1))
        _check(self.lib.w2d_ctqmc_get_steps(self.h, chain, C.byref(n), _p(m), _p(d), _p(sg)), "steps")
        return dict(gtau_mean_step=m[:n.value], gtau_mid_step=d[:n.value], sign_step=sg[:n.value])

    def model_ops_size(self) -> int:
        """Doubles in the eigenbasis operator blocks (shared by all chains)."""
        b = C.c_int64()
        _check(self.lib.w2d_ctqmc_op_table_bytes(self.h, C.byref(b)), "op_table_bytes")
        return b.value // 8

    def orders(self) -> np.ndarray:
        """Operators in each chain's configuration (2 x expansion order)."""
        n = np.zeros(self.nchains, np.int32)
        _check(self.lib.w2d_ctqmc_get_orders(self.h, _p(n), None), "get_orders")
        return n

    def set_abort(self, flag: bool = True):
        """Raise (or clear) the abort word: sweeps stop at the next measurement
        boundary and keep what they measured (the reference's signal handling)."""
        _check(self.lib.w2d_ctqmc_set_abort(self.h, 1 if flag else 0), "set_abort")

    def aborted(self) -> int:
        """``aborted`` of the last warmup/measure: 0, 1 (data kept) or 2 (none)."""
        a = C.c_int()
        _check(self.lib.w2d_ctqmc_aborted(self.h, C.byref(a)), "aborted")
        return a.value

    def moves_log(self, chain: int, n: int) -> np.ndarray:
        rec = np.zeros(n, dtype=MOVE_DT)
        _check(self.lib.w2d_ctqmc_moves_log(self.h, int(chain), int(n), _p(rec)),
               "w2d_ctqmc_moves_log")
        return rec

    def last_kernel_ms(self):
        """(milliseconds, launches) of the last warmup/measure call."""
        ms = C.c_double()
        nl = C.c_int64()
        _check(self.lib.w2d_ctqmc_last_kernel_ms(self.h, C.byref(ms), C.byref(nl)), "last_kernel_ms")
        return ms.value, nl.value

    # -- configurations -----------------------------------------------------------
    def get_config(self, chain: int = 0) -> dict:
        n = C.c_int()
        outer = C.c_int()
        w = np.zeros(5)
        _check(self.lib.w2d_ctqmc_get_config(self.h, chain, 0, C.byref(n), None, None, None, None,
                                             C.byref(outer), _p(w)), "get_config")
        k = n.value
        tau = np.zeros(max(k, 1))
        orb, spin, ca = (np.zeros(max(k, 1), np.int32) for _ in range(3))
        _check(self.lib.w2d_ctqmc_get_config(self.h, chain, k, C.byref(n), _p(tau), _p(orb), _p(spin),
                                             _p(ca), C.byref(outer), _p(w)), "get_config")
        return dict(n=k, taus=tau[:k], orbs=orb[:k], spins=spin[:k], cas=ca[:k],
                    outer_sst=outer.value, trace_log=w[0], trace_sign=w[1], det_log=w[2],
                    det_sign=w[3], cfgsign=w[4])

    def set_config(self, chain: int, taus, orbs, spins, cas, outer_sst: int) -> dict:
        taus = np.ascontiguousarray(taus, np.float64)
        orbs, spins, cas = (np.ascontiguousarray(x, np.int32) for x in (orbs, spins, cas))
        w = np.zeros(5)
        _check(self.lib.w2d_ctqmc_set_config(self.h, chain, len(taus), _p(taus), _p(orbs), _p(spins),
                                             _p(cas), int(outer_sst), _p(w)), "set_config")
        return dict(trace_log=w[0], trace_sign=w[1], det_log=w[2], det_sign=w[3], cfgsign=w[4])

    # ReuseMCConfig for a batch (cthyb_solver.py:24-86, 328-332 keep one
    # configuration per rank): every chain's configuration, written to and read
    # from one .npz, so that the next solve starts each chain where it stopped
    def save_configs(self, path: str):
        cfgs = [self.get_config(c) for c in range(self.nchains)]
        n = np.array([c["n"] for c in cfgs], np.int64)
        cat = (lambda k, dt: np.concatenate([np.asarray(c[k], dt) for c in cfgs]) if n.sum() else
               np.zeros(0, dt))
        np.savez(path, n=n, outer_sst=np.array([c["outer_sst"] for c in cfgs], np.int32),
                 taus=cat("taus", np.float64), orbs=cat("orbs", np.int32), spins=cat("spins", np.int32),
                 cas=cat("cas", np.int32))

    def load_configs(self, path: str) -> list:
        """Set every chain to the configuration saved for it (chains beyond the
        saved ones keep theirs); returns the weights set_config computed."""
        with np.load(path) as z:
            n, outer = z["n"], z["outer_sst"]
            off = np.concatenate([[0], np.cumsum(n)])
            arrs = {k: z[k] for k in ("taus", "orbs", "spins", "cas")}
        out = []
        for c in range(min(self.nchains, len(n))):
            a, b = int(off[c]), int(off[c + 1])
            out.append(self.set_config(c, arrs["taus"][a:b], arrs["orbs"][a:b], arrs["spins"][a:b],
                                       arrs["cas"][a:b], int(outer[c])))
        return out

    def get_minv(self, chain: int, f: int) -> np.ndarray:
        k = C.c_int()
        _check(self.lib.w2d_ctqmc_get_minv(self.h, chain, f, C.byref(k), None), "get_minv")
        out = np.zeros(max(k.value ** 2, 1))
        _check(self.lib.w2d_ctqmc_get_minv(self.h, chain, f, C.byref(k), _p(out)), "get_minv")
        return out[:k.value ** 2].reshape(k.value, k.value, order="F")

    # -- model and results ------------------------------------------------------------
    def model(self) -> dict:
        if self._model is None:
            sizes = np.zeros(self.nsst, np.int32)
            offs = np.zeros(self.nsst, np.int32)
            ev = np.zeros(self.nstates)
            st = np.zeros(self.nstates, np.int32)
            egs = np.zeros(1)
            _check(self.lib.w2d_ctqmc_get_model(self.h, _p(sizes), _p(offs), _p(ev), _p(st), None,
                                                _p(egs)), "get_model")
            vec = np.zeros(int((sizes.astype(np.int64) ** 2).sum()))
            _check(self.lib.w2d_ctqmc_get_model(self.h, None, None, None, None, _p(vec), None),
                   "get_model")
            evecs, o = [], 0
            for n in sizes:
                evecs.append(vec[o:o + n * n].reshape(n, n, order="F"))
                o += n * n
            self._model = dict(sizes=sizes, offsets=offs, evals=ev, states=st, evecs=evecs,
                               egs=float(egs[0]))
        return self._model

    def accumulators(self) -> dict:
        nc, nf = self.nchains, 2 * self.nbands
        g = np.zeros((nc, nf, self.ntau))
        gl = np.zeros((nc, nf, self.nleg))
        dm = np.zeros((nc, self.nstates, self.nstates))
        hi = np.zeros((nc, nf, self.maxhisto + 1))
        oh = np.zeros((nc, self.nsst))
        sc = np.zeros((nc, 14))
        _check(self.lib.w2d_ctqmc_get_accumulators(self.h, _p(g), _p(gl), _p(dm), _p(hi), _p(oh),
                                                   _p(sc)), "get_accumulators")
        # dm rows are column-major per chain
        dm = np.transpose(dm, (0, 2, 1))
        out = dict(gtau=g, gleg=gl, dm=dm, histo=hi, outer_hist=oh, scalars=sc)
        if self.niw:
            out["giw"] = self.giw_sums(0)
        return out

    @property
    def niw(self) -> int:
        """Matsubara frequencies of MeasGiw (0: the handle does not measure G(i w))."""
        n = C.c_int()
        _check(self.lib.w2d_ctqmc_get_giw(self.h, 0, C.byref(n), None), "get_giw")
        return n.value

    def giw_sums(self, ngroups: int = 0) -> np.ndarray:
        """MeasGiw: raw sums of cfgsign M(c,a) exp(i w_n (tau_a - tau_c)) for
        n = 0..Niw-1, [chain or group][flavour][n][re, im]; ngroups = 0 per chain,
        else over the groups of ``group_sums``."""
        n = C.c_int()
        _check(self.lib.w2d_ctqmc_get_giw(self.h, 0, C.byref(n), None), "get_giw")
        if n.value == 0:
            raise CtqmcError(6, "get_giw (the handle was created without MeasGiw)")
        rows = self.nchains if ngroups == 0 else int(ngroups)
        out = np.zeros((rows, 2 * self.nbands, n.value, 2))
        _check(self.lib.w2d_ctqmc_get_giw(self.h, int(ngroups), C.byref(n), _p(out)), "get_giw")
        return out

    def results(self, acc: dict | None = None) -> dict:
        """Per-chain observables normalised as in ctqmc_measure (each chain like
        one reference rank); leading axis = chain.  A chain whose own sign sum
        vanishes gives NaN; ``aggregate`` pools chains instead."""
        if acc is None:
            acc = self.accumulators()
        return normalise(acc, self.nbands, self.beta, self.ntau, self.nleg, self.model())

    def gtau_blocks(self, chain: int = 0) -> np.ndarray:
        """AccumGtau = 2: variances of the means of bins of 2^l measurements of
        chain ``chain``'s raw G(tau), [level][flavour][ntau] (gtau_blocks)."""
        nl = C.c_int()
        _check(self.lib.w2d_ctqmc_get_gtau_blocks(self.h, int(chain), C.byref(nl), None), "gtau_blocks")
        out = np.zeros(nl.value * 2 * self.nbands * self.ntau)
        _check(self.lib.w2d_ctqmc_get_gtau_blocks(self.h, int(chain), C.byref(nl), _p(out)), "gtau_blocks")
        return out.reshape(nl.value, 2 * self.nbands, self.ntau)

    def group_sums(self, ngroups: int, with_dm: bool = True) -> dict:
        """Raw accumulators summed on the GPU over ``ngroups`` contiguous groups
        of chains, leaving out chains stopped on an error; the density matrix
        comes as its full (nstates x nstates) form per group (left out, with
        the occupations that derive from it, when ``with_dm`` is false)."""
        nf, ng = 2 * self.nbands, int(ngroups)
        m = self.model()
        dmsize = int((m["sizes"].astype(np.int64) ** 2).sum())
        g = np.zeros((ng, nf, self.ntau))
        gl = np.zeros((ng, nf, self.nleg))
        db = np.zeros((ng, dmsize)) if with_dm else None
        hi = np.zeros((ng, nf, self.maxhisto + 1))
        oh = np.zeros((ng, self.nsst))
        sc = np.zeros((ng, 14))
        nu = np.zeros(ng, np.int32)
        _check(self.lib.w2d_ctqmc_get_group_sums(self.h, ng, _p(g), _p(gl), _p(db) if with_dm else None,
                                                 _p(hi), _p(oh), _p(sc), _p(nu)), "get_group_sums")
        out = dict(gtau=g, gleg=gl, histo=hi, outer_hist=oh, scalars=sc, nused=nu)
        if self.niw:
            out["giw"] = self.giw_sums(ng)
        if with_dm:
            dm = np.zeros((ng, self.nstates, self.nstates))
            o = 0
            for off, n in zip(m["offsets"], m["sizes"]):
                dm[:, off:off + n, off:off + n] = db[:, o:o + n * n].reshape(ng, n, n).transpose(0, 2, 1)
                o += n * n
            out["dm"] = dm
        return out

    def aggregate(self, ngroups: int = 64) -> dict:
        """Estimates over all chains that completed: the sign-ratio estimate
        (sums of sign-weighted measurements over chains divided by the sum of
        signs, i.e. CTQMC.F90:8663-8771 applied to pooled sums) and its
        jackknife standard error over groups of chains (``<key>_err``).  Unlike
        a mean of per-chain ratios it stays finite where a single chain's sign
        sum vanishes; without a sign problem it agrees with the
        DistributedSample mean (statistics.py:180-209) to O(1/chains)."""
        ok = int((self.chain_status()["status"] == 0).sum())
        ng = max(1, min(int(ngroups), ok))
        sums = self.group_sums(ng)
        hist = self.group_histograms(ng)
        nacc = self.pair_stats(0)["accpair"].shape[1]
        hist["accpair"] = hist["accpair"][:, :, :nacc]
        sums.update(hist)
        sums = {k: v[sums["nused"] > 0] for k, v in sums.items()}
        check_accumulators(sums)
        return pooled_estimate(sums, self.nbands, self.beta, self.ntau, self.nleg, self.model())


    # -- launch machinery ---------------------------------------------------------
    def set_launch(self, graphs: bool = True, order_chains: bool = False, moves_per_launch: int = 0):
        """Sweep launches of a pass captured into one HIP graph; chains ordered
        longest first before every launch; moves per chain and launch (0: ~4000)."""
        _check(self.lib.w2d_ctqmc_set_launch(self.h, int(bool(graphs)), int(bool(order_chains)),
                                             int(moves_per_launch)), "set_launch")

    def engine(self) -> tuple:
        """(trace engine in use, smallest admissible): padded block sizes 1, 4, 16, 128."""
        a, b = C.c_int(), C.c_int()
        _check(self.lib.w2d_ctqmc_get_engine(self.h, C.byref(a), C.byref(b)), "get_engine")
        return a.value, b.value

    def set_engine(self, nm: int):
        _check(self.lib.w2d_ctqmc_set_engine(self.h, int(nm)), "set_engine")

    def autotune(self, moves: int = 2000) -> dict:
        """Time every admissible trace engine over ``moves`` warm-up moves per
        chain and keep the fastest: {engine: moves/s}, and the choice under 'chosen'."""
        nm = C.c_int()
        rates = np.zeros(4)
        _check(self.lib.w2d_ctqmc_autotune(self.h, int(moves), C.byref(nm), _p(rates)), "autotune")
        out = {e: float(r) for e, r in zip((1, 4, 16, 128), rates) if r > 0}
        out["chosen"] = nm.value
        return out

    # -- parameter derivatives (no counterpart in the reference) -----------------
    def set_derivatives(self, deriv, score: bool = True):
        """Derivatives of the create inputs with respect to parameters theta (an
        ``problems.InputDerivatives``: du, dmuimp, dftau per theta, 1..4 of
        them); with ``score`` every later measurement adds the score sums."""
        nt = len(deriv.du)
        n2 = 2 * self.nbands
        du = np.ascontiguousarray(np.stack([np.asarray(d).reshape(n2, n2, n2, n2).ravel(order="F")
                                            for d in deriv.du]))
        dmu = np.ascontiguousarray(np.stack([np.asarray(d).ravel(order="F") for d in deriv.dmuimp]))
        dft = np.ascontiguousarray(np.stack([np.asarray(d).ravel(order="F") for d in deriv.dftau]))
        _check(self.lib.w2d_ctqmc_set_derivatives(self.h, nt, _p(du), _p(dmu), _p(dft), int(bool(score))),
               "set_derivatives")
        self.ntheta = nt
        self._deriv = deriv

    def dlogw(self, chain: int = 0) -> np.ndarray:
        """[theta][0: d ln|det Delta|, 1: d ln|Tr|] of chain ``chain``'s configuration."""
        out = np.zeros(2 * self.ntheta)
        _check(self.lib.w2d_ctqmc_dlogw(self.h, int(chain), _p(out)), "dlogw")
        return out.reshape(self.ntheta, 2)

    def score_sums(self, ngroups: int) -> dict:
        """Score sums of the last measure over ``ngroups`` groups of chains:
        ssum [g][theta], gtau [g][theta][f][tau], dm and dmx [g][theta][state][state]
        (eigenbasis, block-diagonal)."""
        nt, nf, ng = self.ntheta, 2 * self.nbands, int(ngroups)
        m = self.model()
        dmsize = int((m["sizes"].astype(np.int64) ** 2).sum())
        ss = np.zeros((ng, nt))
        g = np.zeros((ng, nt, nf, self.ntau))
        db = np.zeros((ng, nt, dmsize))
        dx = np.zeros((ng, nt, dmsize))
        _check(self.lib.w2d_ctqmc_get_score_sums(self.h, ng, _p(ss), _p(g), _p(db), _p(dx)), "get_score_sums")
        out = dict(ssum=ss, gtau=g)
        for key, blk in (("dm", db), ("dmx", dx)):
            dm = np.zeros((ng, nt, self.nstates, self.nstates))
            o = 0
            for off, n in zip(m["offsets"], m["sizes"]):
                dm[:, :, off:off + n, off:off + n] = blk[:, :, o:o + n * n].reshape(ng, nt, n, n).swapaxes(2, 3)
                o += n * n
            out[key] = dm
        return out

    def score_estimate(self, ngroups: int = 64) -> dict:
        """Observables and their derivatives d<O>/dtheta (score-function
        estimates) pooled over all completed chains, jackknife errors over
        groups of chains; see ``derivatives.score_estimate``."""
        from .derivatives import score_estimate
        ok = int((self.chain_status()["status"] == 0).sum())
        ng = max(1, min(int(ngroups), ok))
        sums = self.group_sums(ng)
        sc = self.score_sums(ng)
        keep = sums["nused"] > 0
        sums = {k: v[keep] for k, v in sums.items()}
        sc = {k: v[keep] for k, v in sc.items()}
        check_accumulators(sums)
        return score_estimate(sums, sc, self.nbands, self.beta, self.ntau, self.model(),
                              self._ftau_diag(), _dftau_diag(self._deriv, self.nbands))

    def _ftau_diag(self) -> np.ndarray:
        return _diag_flavours(self._inputs[1], self.nbands)


def _diag_flavours(ftau, nb: int) -> np.ndarray:
    """F(b, s, b, s, :) of each flavour f = b + N s."""
    ftau = np.asarray(ftau)
    return np.stack([ftau[b, s, b, s] for s in range(2) for b in range(nb)])


def _dftau_diag(deriv, nb: int) -> np.ndarray:
    return np.stack([_diag_flavours(d, nb) for d in deriv.dftau])


def _param_value(params: str, key: str) -> float:
    for line in params.split("\n"):
        if "=" in line and line.split("=", 1)[0].strip().startswith(key):
            return float(line.split("=", 1)[1].split("#")[0].strip())
    raise KeyError(key)


class AccumulationError(RuntimeError):
    pass


def check_accumulators(acc: dict, rtol: float = 1e-9):
    """Guard against lost accumulations: every (annihilator, creator) pair adds
    the same M * cfgsign to one G(tau) bin and, with P_0 = 1, to the l = 0
    Legendre sum (CTQMC.F90:2360-2393), so for every chain (or group) and
    flavour the raw bin sum equals the raw l = 0 sum up to rounding.  Raises
    AccumulationError otherwise."""
    g = np.asarray(acc["gtau"], dtype=float).sum(-1)
    l0 = np.asarray(acc["gleg"], dtype=float)[..., 0]
    scale = np.abs(acc["gtau"]).sum(-1) + np.abs(l0) + 1e-300
    bad = np.abs(g - l0) > rtol * scale
    if bad.any():
        idx = np.argwhere(bad)[0]
        raise AccumulationError("G(tau) bins and l=0 Legendre sum differ at %s: %r vs %r" % (
            tuple(int(i) for i in idx), float(g[tuple(idx)]), float(l0[tuple(idx)])))


def pooled_estimate(sums: dict, nb: int, beta: float, ntau: int, nleg: int, model: dict) -> dict:
    """Normalised observables of the pooled sums over the leading (group) axis
    of ``sums`` and their jackknife errors over that axis."""
    keys = [k for k in ("gtau", "gleg", "dm", "histo", "outer_hist", "scalars", "outer_state", "sign_sst",
                        "sign_state", "contrib_sst", "contrib_state", "accpair", "giw") if k in sums]
    tot = {k: np.asarray(sums[k], dtype=float).sum(0)[None] for k in keys}
    est = normalise(tot, nb, beta, ntau, nleg, model)
    out = {k: np.asarray(v)[0] for k, v in est.items()}
    if "densitymatrix" in out:
        out["rho1"], out["rho2"] = reduced_density_matrices(out["densitymatrix"], nb, model["states"])
    n = len(sums["scalars"])
    if n < 2:
        for k in est:
            out[k + "_err"] = np.full_like(out[k], np.nan, dtype=float)
        return out
    jk = {k: [] for k in est}
    for i in range(n):
        part = {k: tot[k] - np.asarray(sums[k], dtype=float)[i][None] for k in keys}
        r = normalise(part, nb, beta, ntau, nleg, model)
        for k in est:
            jk[k].append(np.asarray(r[k])[0])
    for k in est:
        a = np.array(jk[k])
        out[k + "_err"] = np.sqrt((n - 1) / n * (np.abs(a - a.mean(0)) ** 2).sum(0))
    out["nchains_groups"] = n
    return out


def normalise(acc: dict, nb: int, beta: float, ntau: int, nleg: int, model: dict) -> dict:
    """Normalisation of CTQMC.F90:8663-8771 for every chain of ``acc``."""
    with np.errstate(invalid="ignore", divide="ignore"):
        return _normalise(acc, nb, beta, ntau, nleg, model)


def _normalise(acc, nb, beta, ntau, nleg, model):
    sc = acc["scalars"]
    cnt = sc[:, 1]
    mean_sign = sc[:, 0] / cnt
    zs = -1.0 / (cnt * mean_sign)
    nc = sc.shape[0]

    def fl2(a):   # [chain, f, ...] -> [chain, b, s, ...] with f = b + nb*s
        return a.reshape((nc, 2, nb) + a.shape[2:]).swapaxes(1, 2)

    gtau = fl2(acc["gtau"]) / cnt[:, None, None, None]
    gtau = -gtau / (-mean_sign)[:, None, None, None] * ntau / beta ** 2
    lfac = np.sqrt(2.0 * np.arange(nleg) + 1.0) / beta
    gleg = -fl2(acc["gleg"]) * lfac * zs[:, None, None, None]
    dm_eb = acc["dm"] * zs[:, None, None]
    dm_occ = dm_eb.copy()
    for s, (o, n) in enumerate(zip(model["offsets"], model["sizes"])):
        v = model["evecs"][s]
        dm_occ[:, o:o + n, o:o + n] = v @ dm_eb[:, o:o + n, o:o + n] @ v.T
    diag = np.diagonal(dm_occ, axis1=1, axis2=2)
    states = model["states"]
    nocc = np.array([[(states >> (b + nb * s)) & 1 for s in range(2)] for b in range(nb)])
    occ = np.einsum("ck,ijk,lmk->cijlm", diag, nocc, nocc)
    accr = np.stack([sc[:, 2] / sc[:, 3], sc[:, 4] / sc[:, 5], sc[:, 6] / sc[:, 7],
                     sc[:, 8] / sc[:, 9]], axis=1)
    out = dict(mean_sign=mean_sign, gtau=gtau, gleg=gleg, occ=occ, densitymatrix=dm_occ,
               histo=fl2(acc["histo"]) / cnt[:, None, None, None], acc=accr,
               outer_sst_histo=acc["outer_hist"] / cnt[:, None])
    out.update(normalise_histograms(acc, cnt, nb, beta))
    if "giw" in acc:
        # Giw * zScaling / beta (CTQMC.F90:8816) of the lookup path's -cfgsign M
        # sums: sum / (count x mean sign x beta); n = -Niw..Niw-1, n < 0 the
        # complex conjugates of n' = -n-1
        raw = np.asarray(acc["giw"], dtype=float)
        pos = fl2(raw[..., 0] + 1j * raw[..., 1]) / (cnt * mean_sign * beta)[:, None, None, None]
        out["giw"] = np.concatenate([np.conj(pos[..., ::-1]), pos], axis=-1)
    return out


def _apply_string(nb: int, ops):
    """A product of fermionic operators (applied right to left; (f, dagger),
    f = 2 b + s) on every occupation-basis state: target states, signs and
    the states it does not annihilate.  Bit b + N s holds orbital b, spin s;
    the sign counts the occupied lower bits (Operator.F90:160-172)."""
    st = np.arange(1 << (2 * nb))
    sign = np.ones(st.shape)
    ok = np.ones(st.shape, bool)
    for f, dag in reversed(ops):
        bit = f // 2 + nb * (f % 2)
        occ = (st >> bit) & 1
        ok &= occ != dag
        low = st & ((1 << bit) - 1)
        par = np.zeros(st.shape, np.int64)
        for k in range(bit):
            par += (low >> k) & 1
        sign = np.where(par % 2 == 1, -sign, sign)
        st = st ^ (1 << bit)
    return st, sign, ok


def reduced_density_matrices(dm_occ: np.ndarray, nb: int, states):
    """rho1[f2, f1] = <c+_f2 c_f1> and rho2[f4, f3, f2, f1] =
    <c+_f4 c+_f3 c_f2 c_f1> (f = 2 b + s) from the occupation-basis density
    matrix, as Meas_rho1/rho2_from_Densitymatrix (CTQMC.F90:2936-3060) give
    them: Tr[rho O] = sum_i rho[i, O(i)] sign_i.  Index i of dm_occ is the
    Fock state states[i] (OccbasisMapping)."""
    nf = 2 * nb
    st = np.asarray(states)
    dm = np.zeros((1 << nf, 1 << nf))
    dm[np.ix_(st, st)] = dm_occ
    dm_occ = dm
    idx = np.arange(1 << nf)
    rho1 = np.zeros((nf, nf))
    for f2 in range(nf):
        for f1 in range(nf):
            t, sg, ok = _apply_string(nb, [(f2, 1), (f1, 0)])
            rho1[f2, f1] = np.sum(dm_occ[idx[ok], t[ok]] * sg[ok])
    ann = {(f2, f1): _apply_string(nb, [(f2, 0), (f1, 0)]) for f2 in range(nf) for f1 in range(nf)}
    cre = {(f4, f3): _apply_string(nb, [(f4, 1), (f3, 1)]) for f4 in range(nf) for f3 in range(nf)}
    rho2 = np.zeros((nf,) * 4)
    for (f2, f1), (ta, sa, oka) in ann.items():
        ia = idx[oka]
        if not len(ia):
            continue
        mid, sm = ta[oka], sa[oka]
        for (f4, f3), (tb, sb, okb) in cre.items():
            keep = okb[mid]
            if keep.any():
                rho2[f4, f3, f2, f1] = np.sum(dm_occ[ia[keep], tb[mid[keep]]] * sm[keep] * sb[mid[keep]])
    return rho1, rho2


def normalise_histograms(acc: dict, cnt, nb: int, beta: float) -> dict:
    """Post-processing of the outer-state histograms, AccPair and the
    step-resolved arrays (CTQMC.F90:8741-8764) for those present in ``acc``:
    signs per visit, visits and trace contributions per measurement, the step
    sums divided by beta^2; AccPair stays a count."""
    with np.errstate(invalid="ignore", divide="ignore"):
        return _normalise_histograms(acc, cnt, nb, beta)


def _normalise_histograms(acc, cnt, nb, beta):
    out = {}
    c = np.asarray(cnt, dtype=float)[:, None]
    nc = c.shape[0]
    if "outer_state" in acc:
        out["sign_sst"] = acc["sign_sst"] / acc["outer_hist"]
        out["sign_state"] = acc["sign_state"] / acc["outer_state"]
        out["outer_state_histo"] = acc["outer_state"] / c
        out["contrib_sst"] = acc["contrib_sst"] / c
        out["contrib_state"] = acc["contrib_state"] / c
    if "accpair" in acc:
        a = np.asarray(acc["accpair"])
        out["accpair"] = a.reshape((nc, 2, nb) + a.shape[2:]).swapaxes(1, 2)
    for k in ("gtau_mean_step", "gtau_mid_step"):
        if k in acc:
            a = np.asarray(acc[k]) / beta ** 2                  # [chain, step, f]
            out[k] = a.reshape(a.shape[:2] + (2, nb)).transpose(0, 3, 2, 1)   # [chain, b, s, step]
    if "sign_step" in acc:
        out["sign_step"] = np.asarray(acc["sign_step"])
    return out
