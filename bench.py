"""Benchmark: CT-HYB Metropolis moves/s and G(tau) 1/err^2/s on one MI355X,
batched chains vs the plain single-chain path.

Workload (default, c2): 2-orbital Kanamori impurity (U=4, J=1), half filling,
beta=50, semicircular baths of half-bandwidth 2, 16 local states in (N,Sz) blocks
of up to 4 states, 1024 batched Markov chains, NCorr=100, TaudiffMax=5.  Other
workloads: c1 (1-orbital Anderson), c3 (3-orbital Kanamori, beta=100, 4096
chains, blocks up to 9 x 9), c4 (5-orbital full-Coulomb impurity, F0=4, J=0.8,
filling 6, beta=40, 1024 states in blocks up to 100 x 100 on the FP64 matrix
cores, 2048 chains), c5 (c3 with score-function gradients).  A "step" is 40
sweeps of every chain for c1/c2 (one launch of the persistent sweep kernel), 4
for c3 and 1 for c4 and c5, a sweep being NCorr Metropolis moves followed by the
Z-space measurements (G(tau) bins + Legendre, beta/2 density matrix, sign,
histograms).

Before the timed region the chains are thermalised until the mean expansion
order over all chains stops growing (whatever --warmup says; --warmup is the
minimum number of sweeps), and the mean order is reported before and after the
timed region.  The plain single-chain path is one chain of the same engine,
started from a thermalised configuration and measured with the binning
accumulator (AccumGtau = 2); its G(tau) error is the binning plateau.  The CPU
baseline is the reference Fortran core built from its sources (oracle/_ref),
one process per host core, each started from a thermalised configuration
(set_mc_config, as ReuseMCConfig does) under a fixed measurement budget.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3|c4|c5]
                       [--chains N] [--no-cpu] [--no-plain] [--therm-s S]

Prints one JSON line with a roofline object for the sweep kernel (FP64 MFMA for
blocks above 16 states, HBM otherwise) and the reference CPU path as cpu_baseline.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFS = 78.6       # MI355X dense FP64 (vector and matrix)
METRIC = "MC moves/s and G(τ) 1/err²/s on 1 GPU, batched chains vs plain path"

WORKLOADS = {
    # name: (builder, builder kwargs, TaudiffMax, chains, initial capacity, reference Nmeas per process)
    "c1": ("anderson_1orb", dict(beta=10.0), 2.0, 1024, 128, 20000),
    "c2": ("kanamori_2orb", dict(beta=50.0), 5.0, 1024, 128, 20000),
    "c3": ("kanamori_3orb", dict(beta=100.0), 8.0, 4096, 384, 300),
    "c4": ("coulomb_5orb", dict(beta=40.0), 5.0, 2048, 64, 12),
}
# sweeps (NCorr moves + one measurement per chain) in one step: one launch of
# the sweep kernel (40 sweeps = 4000 moves per chain) for the small-block
# workloads, fewer where a sweep already takes tens of milliseconds
SWEEPS_PER_STEP = {"c1": 40, "c2": 40, "c3": 4, "c4": 1}


def make_problem(cfg, ncorr, **extra):
    from w2dynamics_amd import problems as P
    name, kw, tdm = WORKLOADS[cfg][:3]
    return getattr(P, name)(TaudiffMax=tdm, NCorr=ncorr, **kw, **extra), tdm


def gtau_inv_err2(res):
    """1 / <stderr^2(G(tau))> over flavours and tau bins, stderr across samples."""
    g = np.asarray(res["gtau"])               # [samples, b, s, ntau]
    n = g.shape[0]
    m = g.mean(0)
    se2 = ((g - m) ** 2).sum(0) / (n * (n - 1))
    if not float(se2.mean()) > 0.0:
        print("G(tau) samples carry no spread; 1/err^2 undefined", file=sys.stderr)
        return float("nan")
    return 1.0 / float(se2.mean())


def ratio_inv_err2(graw, sgn, ntau, beta):
    """1 / <stderr^2(G(tau))> of the sign-ratio estimate G = sum G_raw / sum sign
    (the normalisation of CTQMC.F90:8663-8771 applied to pooled sums), by the
    jackknife over the leading axis (groups of chains).  Without a sign problem
    it is the plain standard error of the group means; with one it stays finite
    where a single group's sign sum vanishes."""
    g = np.asarray(graw, dtype=float).reshape(len(sgn), -1)
    s = np.asarray(sgn, dtype=float)
    n = g.shape[0]
    if n < 2:
        return float("nan")
    jk = (g.sum(0)[None, :] - g) / (s.sum() - s)[:, None] * ntau / beta ** 2
    se2 = (n - 1) / n * ((jk - jk.mean(0)) ** 2).sum(0)
    if not (np.all(np.isfinite(se2)) and float(se2.mean()) > 0.0):
        print("G(tau) jackknife carries no spread or no sign; 1/err^2 undefined", file=sys.stderr)
        return float("nan")
    return 1.0 / float(se2.mean())


def plateau_inv_err2(blocks, nmeas, mean_sign, ntau, beta, min_bins=16):
    """Single chain: binning analysis of the raw G(tau) (gtau_blocks of the
    autocorracc accumulator, Accumulators.F90:283-362).  Level l holds the
    variance of the means of bins of 2^l measurements; stderr^2 of the mean at
    level l is var_l / n_l with n_l = nmeas >> l.  The plateau is the largest
    <stderr^2> over the levels with at least `min_bins` bins (conservative);
    the normalisation of G (1/|mean sign| * ntau / beta^2) is applied."""
    curve = []
    for lvl in range(blocks.shape[0]):
        nl = nmeas >> lvl
        if nl < min_bins:
            break
        se2 = blocks[lvl] / nl * (ntau / beta ** 2 / abs(mean_sign)) ** 2
        curve.append(float(np.mean(se2)))
    if not curve or not max(curve) > 0:
        return float("nan"), curve
    return 1.0 / max(curve), curve


def thermalise(gpu, ncorr, min_sweeps, chunk, budget_s, tol=0.015):
    """Warm up in chunks of `chunk` sweeps until the mean order over all chains
    has stopped growing (the mean of the last two chunks within `tol` of the
    two before) and at least `min_sweeps` sweeps are done, or `budget_s` is
    spent.  Returns the trajectory of (sweeps, mean order, seconds)."""
    traj = [(0, float(gpu.orders().mean()), 0.0)]
    t0 = time.perf_counter()
    sweeps = 0
    while True:
        gpu.warmup(chunk * ncorr)
        sweeps += chunk
        traj.append((sweeps, float(gpu.orders().mean()), time.perf_counter() - t0))
        print("thermalisation: %d sweeps, mean order %.2f, %.1f s" % traj[-1], file=sys.stderr,
              flush=True)
        m = [x[1] for x in traj[1:]]
        if sweeps >= min_sweeps and len(m) >= 4 and \
                abs(np.mean(m[-2:]) - np.mean(m[-4:-2])) <= tol * max(1.0, np.mean(m[-2:])):
            break
        if time.perf_counter() - t0 > budget_s:
            print("thermalisation budget of %.0f s spent before the order settled" % budget_s,
                  file=sys.stderr)
            break
    return traj


def run_gpu(cfg, nchains, capacity, min_sweeps, steps, ncorr, seed, therm_s, launch=None, engine=0, niw=0):
    import torch
    from w2dynamics_amd.ctqmc import CtqmcGpu, check_accumulators
    prob, tdm = make_problem(cfg, ncorr, **({"MeasGiw": 1, "Niw": niw} if niw else {}))
    gpu = CtqmcGpu.from_problem(prob, nchains=nchains, seed=seed, capacity=capacity, device=0)
    gpu.set_launch(**(launch or {}))
    if engine > 0:
        gpu.set_engine(engine)
    gpu.set_taudiffmax(tdm)
    chunk = 5 if cfg in ("c3", "c4") else 20
    traj = thermalise(gpu, ncorr, min_sweeps, chunk, therm_s)
    tuned = gpu.autotune(2 * ncorr) if engine < 0 else None
    order_before = float(gpu.orders().mean())
    print("timed measurement: %d sweeps" % steps, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.measure(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, nlaunch = gpu.last_kernel_ms()
    order_after = float(gpu.orders().mean())
    ng = min(64, nchains)
    sums = gpu.group_sums(ng, with_dm=False)
    check_accumulators(sums)       # raises if any G(tau) accumulation was lost
    sc = sums["scalars"]
    q = prob.qmc()
    giw_flops = None
    if niw:
        # G(i w) kernel work per launch, from the configurations at the end:
        # per flavour k^2 complex-by-real products (4 flops) and k complex
        # products (8 flops) for every frequency
        fl = 0.0
        for c in range(nchains):
            x = gpu.get_config(c)
            for b in range(prob.nbands):
                for s in range(2):
                    k = int(((x["orbs"] == b) & (x["spins"] == s)).sum()) // 2
                    fl += (4.0 * k * k + 8.0 * k) * niw
        giw_flops = fl
        assert np.isfinite(gpu.results()["giw"]).all()
    out = dict(wall=wall, kernel_ms=kms, launches=nlaunch, giw_flops=giw_flops,
               inv_err2=ratio_inv_err2(sums["gtau"], sc[:, 0], int(q["Ntau"]), prob.beta),
               moves=float(sc[:, 10].sum()), alg_flops=float(sc[:, 12].sum()),
               alg_bytes=float(sc[:, 13].sum()), mean_sign=float(sc[:, 0].sum() / sc[:, 1].sum()),
               order_before=order_before, order_after=order_after, traj=traj,
               status=gpu.chain_status(), nops_bytes=8.0 * gpu.model_ops_size(),
               configs=[gpu.get_config(c) for c in range(min(nchains, 64))], label=prob.label,
               beta=prob.beta, tdm=tdm, engine=gpu.engine(), tuned=tuned)
    gpu.close()
    return out


def run_plain(cfg, capacity, ncorr, seed, start_cfg, budget_s=25.0, engine=0):
    """Single chain from a thermalised configuration; binning-plateau error."""
    import torch
    from w2dynamics_amd.ctqmc import CtqmcGpu
    # a short timed run sizes the measurement to the budget
    prob, tdm = make_problem(cfg, ncorr, AccumGtau=2, Nmeas=4)
    gpu = CtqmcGpu.from_problem(prob, nchains=1, seed=seed, capacity=capacity, device=0)
    if engine > 0:
        gpu.set_engine(engine)
    gpu.set_taudiffmax(tdm)
    c = start_cfg
    gpu.set_config(0, c["taus"], c["orbs"], c["spins"], c["cas"], c["outer_sst"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.measure(4)
    torch.cuda.synchronize()
    rate = 4 * ncorr / (time.perf_counter() - t0)
    gpu.close()
    nmeas = int(min(8192, max(64, budget_s * rate / ncorr)))
    prob, tdm = make_problem(cfg, ncorr, AccumGtau=2, Nmeas=nmeas)
    gpu = CtqmcGpu.from_problem(prob, nchains=1, seed=seed, capacity=capacity, device=0)
    if engine > 0:
        gpu.set_engine(engine)
    gpu.set_taudiffmax(tdm)
    gpu.set_config(0, c["taus"], c["orbs"], c["spins"], c["cas"], c["outer_sst"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.measure(nmeas)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sc = gpu.group_sums(1, with_dm=False)["scalars"][0]
    blocks = gpu.gtau_blocks(0)
    q = prob.qmc()
    inv, curve = plateau_inv_err2(blocks, nmeas, sc[0] / sc[1], int(q["Ntau"]), prob.beta)
    order = float(gpu.orders().mean())
    gpu.close()
    return dict(wall=wall, moves=float(nmeas * ncorr), inv_err2=inv, nmeas=nmeas, curve=curve,
                order=order)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def write_configs(path, cfg):
    """configs.bin of the reference driver: int32 count, int32 n, int32 outer
    superstate (0-based), n taus, then orbitals, spins (1-based) and c/a codes."""
    with open(path, "wb") as f:
        np.array([1, cfg["n"], cfg["outer_sst"]], "<i4").tofile(f)
        np.asarray(cfg["taus"], "<f8").tofile(f)
        (np.asarray(cfg["orbs"]) + 1).astype("<i4").tofile(f)
        (np.asarray(cfg["spins"]) + 1).astype("<i4").tofile(f)
        np.asarray(cfg["cas"]).astype("<i4").tofile(f)


class CpuBaseline:
    """The reference CPU path (oracle/_ref/ref_driver, built from the reference
    sources) on this host's cores: one process per core (seeds NSeed + i, like
    MPI ranks), each started from a thermalised configuration of a GPU chain
    and measuring a fixed number of sweeps.  The processes start before the
    GPU run (the reference's model setup takes minutes for 1024 states) and
    wait on standard input until their configurations are written.  Falls back
    to the C oracle (a single-threaded port) if the reference build is absent."""

    def __init__(self, cfg, ncorr, seed):
        self.cfg, self.ncorr, self.seed = cfg, ncorr, seed
        self.nmeas = WORKLOADS[cfg][5]
        self.prob, self.tdm = make_problem(cfg, ncorr, Nmeas=self.nmeas)
        self.procs, self.dirs = [], []
        drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
        try:
            ncpu = len(os.sched_getaffinity(0))
        except AttributeError:
            ncpu = os.cpu_count() or 1
        # one core stays with the GPU process
        self.nproc = max(1, min(15, ncpu - 1))
        if not os.path.exists(drv):
            return
        try:
            sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
            import make_golden as MG
            for i in range(self.nproc):
                d = tempfile.mkdtemp(prefix="w2dref")
                MG.write_inputs(d, self.prob, "runcfg", seed + i, 0, 0)
                self.dirs.append(d)
            # one core per process, like one MPI rank per core: no BLAS/OpenMP
            # threads of their own (many threads per process oversubscribe the
            # cores and stall the model setup)
            env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1")
            self.procs = [subprocess.Popen([drv, d], stdin=subprocess.PIPE, stdout=subprocess.DEVNULL,
                                           stderr=subprocess.DEVNULL, env=env) for d in self.dirs]
        except Exception as e:  # reported, then the port is timed instead
            print("reference CPU run could not start (%s)" % e, file=sys.stderr)
            self.procs = []

    def finish(self, configs):
        if self.procs:
            try:
                return self._finish_reference(configs)
            except Exception as e:  # reported, then fall back to the port
                print("reference CPU run failed (%s); timing the C oracle instead" % e, file=sys.stderr)
        return self._port(configs)

    def _finish_reference(self, configs):
        n = min(len(self.procs), len(configs))
        for i, (p, d) in enumerate(zip(self.procs, self.dirs)):
            if i < n:
                write_configs(os.path.join(d, "configs.bin"), configs[i])
                p.stdin.write(b"go\n")
            p.stdin.close()
        t0 = time.perf_counter()
        rcs = [p.wait() for p in self.procs]
        wall = time.perf_counter() - t0
        if not all(rc == 0 for rc in rcs[:n]):
            raise RuntimeError("reference runs returned %s" % rcs)
        nb, ntau = self.prob.nbands, int(self.prob.qmc()["Ntau"])
        g, tsim, orders = [], [], []
        for d in self.dirs[:n]:
            raw = np.fromfile(os.path.join(d, "results.bin"), dtype="<f8")
            g.append(raw[1:1 + nb * 2 * ntau].reshape((nb, 2, ntau), order="F"))
            tw, ts, nop = open(os.path.join(d, "timing.txt")).read().split()
            tsim.append(float(ts))
            orders.append(int(nop))
        moves = self.nmeas * self.ncorr
        # the processes sample concurrently: the aggregate rate is the sum of
        # each one's moves over its own sampling time (the model setup is not
        # sampling)
        value = float(sum(moves / t for t in tsim))
        return dict(value=value, unit="moves/s", cores=n, kind="reference", cpu=cpu_model(),
                    sample="%d processes x %d measured sweeps of %d moves from thermalised configurations "
                           "(mean order %.0f operators), %s" % (
                               n, self.nmeas, self.ncorr, np.mean([c["n"] for c in configs[:n]]),
                               self.prob.label),
                    gtau_inv_err2_per_s=gtau_inv_err2({"gtau": np.array(g)}) / max(tsim),
                    sampling_s=[round(t, 3) for t in tsim], wall_after_release_s=wall,
                    final_mean_order=float(np.mean(orders)))

    def _port(self, configs):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import OracleChain
        o = OracleChain.from_problem(self.prob, self.seed)
        o.set_taudiffmax(self.tdm)
        c = configs[0]
        o.set_config(c["taus"], c["orbs"], c["spins"], c["cas"], c["outer_sst"])
        t0 = time.perf_counter()
        o.measure(self.nmeas, self.ncorr)
        wall = time.perf_counter() - t0
        return dict(value=self.nmeas * self.ncorr / wall, unit="moves/s", cores=1, kind="port", cpu=cpu_model(),
                    sample="1 chain x %d sweeps of %d moves from a thermalised configuration, %s" % (
                        self.nmeas, self.ncorr, self.prob.label), wall_s=wall)

    def cancel(self):
        for p in self.procs:
            try:
                p.kill()
            except OSError:
                pass


def copy_probe(mib: int):
    """Plain device-to-device copies of known size: the FETCH_SIZE / WRITE_SIZE
    counters of these dispatches calibrate the counters of the sweep kernel."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                      "dispatches": 5}))


def traffic_per_move(workload: str):
    """HBM bytes per attempted move of the sweep kernel measured with rocprofv3
    --pmc (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel) on this workload
    at equilibrium (profiles/sweep_kernel_traffic_<workload>.txt), if recorded."""
    path = os.path.join(ROOT, "profiles", "sweep_kernel_traffic_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_move:"):
                return float(line.split(":", 1)[1])
    return None


def run_c5(args):
    """Differentiation workload: the 3-orbital Kanamori model (beta=100) as a
    function of theta = (U, J, eps_d, V), 4096 chains with score sums at every
    measurement; score-function estimates of d<n_up n_dn>/dU and dG(tau)/dJ
    against central differences with common random numbers (same seeds, the
    chains' thermalised configurations as the common start) at delta = 1e-3."""
    import torch
    from w2dynamics_amd import problems as P
    from w2dynamics_amd.ctqmc import CtqmcGpu
    from w2dynamics_amd.derivatives import finite_difference_crn
    theta = np.array([3.5, 0.6, 0.0, 1.0])
    ncorr, nchains, seed = args.ncorr, args.chains or 4096, args.seed

    def builder(th):
        return P.kanamori_3orb_theta(th, beta=100.0, TaudiffMax=8.0, NCorr=ncorr, Nmeas=args.steps)

    prob, deriv = builder(theta)
    gpu = CtqmcGpu.from_problem(prob, nchains=nchains, seed=seed, capacity=WORKLOADS["c3"][4], device=0)
    gpu.set_taudiffmax(8.0)
    traj = thermalise(gpu, ncorr, args.warmup, 5, args.therm_s)
    # the derivative pass runs on the engines for blocks up to 16 states (the
    # model's default here); no autotuning
    tuned = None
    gpu.set_derivatives(deriv)
    configs = [gpu.get_config(c) for c in range(nchains)]
    print("timed measurement with score sums: %d sweeps" % args.steps, file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.measure(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    est = gpu.score_estimate(64)
    sc = gpu.group_sums(1, with_dm=False)["scalars"][0]
    engine = gpu.engine()[0]
    gpu.close()
    checks = {}
    for t, key in ((0, "double_occ"), (1, "gtau")):
        print("finite differences with common random numbers: %s" % P.KANAMORI_THETA[t], file=sys.stderr, flush=True)
        fd = finite_difference_crn(builder, theta, t, 1e-3, configs, nchains, seed, args.steps,
                                   taudiffmax=8.0, capacity=WORKLOADS["c3"][4], engine=engine)
        s_, se = np.asarray(est["d" + key][t]), np.asarray(est["d" + key + "_err"][t])
        z = np.abs(s_ - fd[key]) / np.sqrt(se ** 2 + fd[key + "_err"] ** 2 + 1e-300)
        checks["d%s/d%s" % (key, P.KANAMORI_THETA[t])] = dict(
            score=np.ravel(s_)[:6].tolist(), score_err=np.ravel(se)[:6].tolist(),
            fd=np.ravel(fd[key])[:6].tolist(), fd_err=np.ravel(fd[key + "_err"])[:6].tolist(),
            within_2sigma=float(np.mean(z <= 2.0)), max_sigma=float(z.max()), entries=int(z.size))
    line = {
        "metric": "score-function gradients: MC moves/s with d ln|w|/dtheta sums at every measurement",
        "value": float(sc[10]) / wall, "unit": "moves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (semicircular-bath 3-orbital Kanamori model, seeded crystal field)",
        "config": {"workload": "c5: 3orb-kanamori beta=100 as a function of (U, J, eps_d, V), %d chains, "
                               "NCorr=%d" % (nchains, ncorr), "trace_engine": engine,
                   "autotune_moves_per_s": tuned, "thermalisation": [[a, round(b, 2), round(c, 1)] for a, b, c in traj]},
        "mean_sign": float(sc[0] / sc[1]),
        "dlogw_mean": np.asarray(est["dlogw_mean"]).tolist(),
        "double_occ": np.asarray(est["double_occ"]).tolist(),
        "ddouble_occ_dU": np.asarray(est["ddouble_occ"][0]).tolist(),
        "ddouble_occ_dU_err": np.asarray(est["ddouble_occ_err"][0]).tolist(),
        "crn_finite_differences_delta_1e-3": checks,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10, help="minimum thermalisation sweeps")
    ap.add_argument("--config", default="c2", choices=sorted(WORKLOADS) + ["c5"])
    ap.add_argument("--chains", type=int, default=0)
    ap.add_argument("--ncorr", type=int, default=100)
    ap.add_argument("--seed", type=int, default=43890)
    ap.add_argument("--therm-s", type=float, default=240.0, help="thermalisation time budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="plain stream launches instead of one HIP graph per pass")
    ap.add_argument("--order", action="store_true", help="order the chains longest first before every launch")
    ap.add_argument("--engine", type=int, default=-1, help="trace engine (padded block size); 0: the model's smallest, "
                    "-1: autotune")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--c4-anyway", action="store_true", help="run the 5-orbital workload although it is switched off")
    ap.add_argument("--niw", type=int, default=0,
                    help="also measure G(i w_n) (MeasGiw) on this many frequencies: measure launches end on every "
                         "measurement and the G(iw) kernel follows each; the roofline is then left out (the "
                         "launches time both kernels)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 5 device copies of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the CT-HYB chains of one handle live on one device "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    cfg = args.config
    if cfg == "c5":
        run_c5(args)
        return
    if cfg == "c4" and not args.c4_anyway:
        # the 5-orbital thermalisation at 2048 chains ended in GPU memory faults
        # (twice at the same point of the seeded run) and its cause is not
        # shown: it is not run unless asked for explicitly
        print("bench.py: --config c4 is switched off (its thermalisation faulted the GPU; cause not "
              "shown); pass --c4-anyway to run it", file=sys.stderr)
        sys.exit(2)
    nchains = args.chains or WORKLOADS[cfg][3]
    capacity = WORKLOADS[cfg][4]
    seed = args.seed

    # the reference CPU processes set their model up while the GPU runs
    cpu_run = None if args.no_cpu else CpuBaseline(cfg, args.ncorr, seed)
    print("batched run: %s, %d chains" % (cfg, nchains), file=sys.stderr, flush=True)
    try:
        b = run_gpu(cfg, nchains, capacity, args.warmup, args.steps * SWEEPS_PER_STEP[cfg], args.ncorr, seed,
                    args.therm_s,
                    launch=dict(graphs=not args.no_graphs, order_chains=args.order), engine=args.engine,
                    niw=args.niw)
    except BaseException:
        if cpu_run:
            cpu_run.cancel()
        raise
    value = b["moves"] / b["wall"]
    inv_err2_s = b["inv_err2"] / b["wall"]
    kernel_s = b["kernel_ms"] / 1e3
    launches = max(1, b["launches"])
    # algorithmic bytes: state vectors and matrices counted per chain by the
    # kernel, operator blocks (shared by all chains) once per launch
    alg_bytes = b["alg_bytes"] + b["nops_bytes"] * launches
    ach_gbs = alg_bytes / kernel_s / 1e9
    ach_tfs = b["alg_flops"] / kernel_s / 1e12
    mfma = cfg == "c4"
    tpm = traffic_per_move(cfg)
    traffic = tpm * b["moves"] / launches if tpm else None
    roof = {"bound": "mfma" if mfma else "hbm",
            "achieved": ach_tfs if mfma else ach_gbs,
            "peak": FP64_PEAK_TFS if mfma else HBM_PEAK_GBS,
            "unit": "TFLOP/s" if mfma else "GB/s",
            "frac": (ach_tfs / FP64_PEAK_TFS) if mfma else (ach_gbs / HBM_PEAK_GBS),
            "traffic": traffic,
            "kernel": "sweep_kernel", "launches": b["launches"], "avg_launch_ms": b["kernel_ms"] / launches,
            "alg_flops_per_launch": b["alg_flops"] / launches, "alg_bytes_per_launch": alg_bytes / launches,
            "fp64_tflops": ach_tfs, "fp64_frac": ach_tfs / FP64_PEAK_TFS,
            "hbm_gbs": ach_gbs, "hbm_frac": ach_gbs / HBM_PEAK_GBS}
    meas_giw = None
    if args.niw:
        meas_giw = {"niw": args.niw, "launches": b["launches"], "kernel": "giw_kernel",
                    "alg_flops_per_launch": b["giw_flops"],
                    "note": "the timed region and the launch times include the G(iw) kernel after every "
                            "measurement; per-kernel times come from rocprofv3"}
        roof = None

    plain = None
    if not args.no_plain:
        print("plain single-chain run", file=sys.stderr, flush=True)
        p = run_plain(cfg, capacity, args.ncorr, seed + 1000003, b["configs"][0], engine=b["engine"][0])
        plain = dict(moves_per_s=p["moves"] / p["wall"], gtau_inv_err2_per_s=p["inv_err2"] / p["wall"],
                     sweeps=p["nmeas"], wall_s=p["wall"], mean_order=p["order"],
                     error="binning plateau (AccumGtau=2): largest <stderr^2> over levels with >= 16 bins",
                     plateau_curve=p["curve"])

    cpu = None
    if cpu_run:
        print("CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_run.finish(b["configs"])

    st = b["status"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (semicircular-bath impurity model, seeded crystal field)",
        "config": {"workload": "%s: %s, beta=%g, %d batched chains, NCorr=%d, TaudiffMax=%g" % (
            cfg, b["label"], b["beta"], nchains, args.ncorr, b["tdm"]),
            "chains": nchains, "ncorr": args.ncorr, "sweeps_per_step": SWEEPS_PER_STEP[cfg],
            "mean_order_ops_before": b["order_before"], "mean_order_ops_after": b["order_after"],
            "thermalisation": [[s, round(o, 2), round(t, 1)] for s, o, t in b["traj"]],
            "capacity": st["capacity"], "max_order_per_flavour": int(st["kfmax"].max()),
            "chains_failed": int((st["status"] != 0).sum()),
            "trace_engine": b["engine"][0], "autotune_moves_per_s": b["tuned"],
            "launch": {"hip_graph": not args.no_graphs, "chains_longest_first": args.order}},
        "gtau_inv_err2_per_s": inv_err2_s,
        "mean_sign": b["mean_sign"],
        "plain_single_chain": plain,
        "batched_vs_plain_inv_err2_ratio": (inv_err2_s / plain["gtau_inv_err2_per_s"]) if plain else None,
        "batched_vs_plain_moves_ratio": (value / plain["moves_per_s"]) if plain else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if meas_giw:
        line["meas_giw"] = meas_giw
    print(json.dumps(line))


if __name__ == "__main__":
    main()
