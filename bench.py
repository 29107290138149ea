"""Benchmark: CT-HYB Metropolis moves/s and G(tau) 1/err^2/s on one MI355X,
batched chains vs the plain single-chain path.

Workload (default): 2-orbital Kanamori impurity, U=4, J=1, U'=U-2J with spin flip
and pair hopping, half filling, beta=50, semicircular baths of half-bandwidth 2,
16 local states in (N,Sz) blocks, FP64; 1024 batched Markov chains.  A "step" is
one sweep: NCorr Metropolis moves of every chain followed by the Z-space
measurements (G(tau) bins + Legendre, beta/2 density matrix, sign, histograms).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--chains 1024]
                       [--config c2|c1|c3|c4] [--no-cpu] [--no-plain]

Prints one JSON line (see the "metric" keys) with a roofline object for the
sweep kernel and the reference CPU path timed on this host as cpu_baseline.
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

WORKLOADS = {
    # name: (builder kwargs, TaudiffMax, chains, capacity)
    "c1": ("anderson_1orb", dict(beta=10.0), 2.0, 1024, 128),
    "c2": ("kanamori_2orb", dict(beta=50.0), 5.0, 1024, 128),
    "c3": ("kanamori_3orb", dict(beta=100.0), 8.0, 4096, 384),
    "c4": ("coulomb_5orb", dict(beta=40.0), 5.0, 2048, 64),
}


def make_problem(cfg, ncorr):
    from w2dynamics_amd import problems as P
    name, kw, tdm, _, _ = WORKLOADS[cfg]
    return getattr(P, name)(TaudiffMax=tdm, NCorr=ncorr, **kw), tdm


def gtau_inv_err2(res):
    """1 / <stderr^2(G(tau))> over flavours and tau bins, stderr across samples."""
    g = np.asarray(res["gtau"])               # [samples, b, s, ntau]
    n = g.shape[0]
    m = g.mean(0)
    se2 = ((g - m) ** 2).sum(0) / (n * (n - 1))
    if not float(se2.mean()) > 0.0:
        print("G(tau) samples carry no spread (sum |G| = %g); 1/err^2 undefined" % float(np.abs(g).sum()),
              file=sys.stderr)
        return float("nan")
    return 1.0 / float(se2.mean())


def ratio_inv_err2(graw, sgn, ntau, beta):
    """1 / <stderr^2(G(tau))> of the sign-ratio estimate G = sum G_raw / sum sign
    (the normalisation of CTQMC.F90:8663-8771 applied to pooled sums), by the
    jackknife over the leading axis (blocks of chains or of measurements).
    Without a sign problem it is the plain standard error of the block means;
    with one it stays finite where a single block's sign sum vanishes."""
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


def gtau_inv_err2_chains(acc, ntau, beta, nblocks=64):
    """Batched run: the chains pooled into `nblocks` contiguous blocks."""
    g = np.asarray(acc["gtau"], dtype=float)            # [chain, flavour, ntau] raw sums
    sg = np.asarray(acc["scalars"], dtype=float)[:, 0]  # sum of signs per chain
    nb = min(nblocks, g.shape[0])
    blocks = np.array_split(np.arange(g.shape[0]), nb)
    return ratio_inv_err2(np.stack([g[i].sum(0) for i in blocks]), np.array([sg[i].sum() for i in blocks]),
                          ntau, beta)


def run_gpu(prob, tdm, nchains, capacity, warm_sweeps, steps, ncorr, seed):
    import torch
    from w2dynamics_amd.ctqmc import CtqmcGpu
    gpu = CtqmcGpu.from_problem(prob, nchains=nchains, seed=seed, capacity=capacity, device=0)
    gpu.set_taudiffmax(tdm)
    gpu.warmup(warm_sweeps * ncorr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.measure(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, nlaunch = gpu.last_kernel_ms()
    acc = gpu.accumulators()
    res = gpu.results(acc)
    out = dict(wall=wall, kernel_ms=kms, launches=nlaunch,
               inv_err2=gtau_inv_err2_chains(acc, prob.qmc()["Ntau"], prob.beta),
               moves=float(acc["scalars"][:, 10].sum()),
               alg_flops=float(acc["scalars"][:, 12].sum()),
               alg_bytes=float(acc["scalars"][:, 13].sum()), res=res,
               order=float(np.mean([gpu.get_config(c)["n"] for c in range(min(nchains, 8))])))
    gpu.close()
    return out


def run_plain(prob, tdm, capacity, warm_sweeps, steps, ncorr, seed, blocks=8):
    """Single chain: the same engine with one chain; error from block means."""
    import torch
    from w2dynamics_amd.ctqmc import CtqmcGpu
    gpu = CtqmcGpu.from_problem(prob, nchains=1, seed=seed, capacity=capacity, device=0)
    gpu.set_taudiffmax(tdm)
    gpu.warmup(warm_sweeps * ncorr)
    per = max(1, steps // blocks)
    g, sg = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(blocks):
        gpu.measure(per)
        acc = gpu.accumulators()
        g.append(np.asarray(acc["gtau"][0], dtype=float))
        sg.append(float(acc["scalars"][0, 0]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu.close()
    moves = blocks * per * ncorr
    return dict(wall=wall, moves=moves,
                inv_err2=ratio_inv_err2(np.array(g), np.array(sg), prob.qmc()["Ntau"], prob.beta))


def cpu_baseline(prob, tdm, seed, nwarm, nmeas, ncorr):
    """The reference CPU path (oracle/_ref/ref_driver, built from the reference
    sources) timed on this host: one process per core, seeds NSeed + i, like MPI
    ranks; falls back to the C oracle (a single-threaded port) if the reference
    build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    nproc = max(1, min(16, ncpu))
    p = prob
    p.params = dict(p.params, Nmeas=nmeas, NCorr=ncorr)
    if os.path.exists(drv):
        try:
            import make_golden as MG
            dirs = []
            for i in range(nproc):
                d = tempfile.mkdtemp(prefix="w2dref")
                MG.write_inputs(d, p, "run", seed + i, nwarm, 0)
                dirs.append(d)
            t0 = time.perf_counter()
            procs = [subprocess.Popen([drv, d], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                     for d in dirs]
            rcs = [q.wait() for q in procs]
            wall = time.perf_counter() - t0
            if all(rc == 0 for rc in rcs):
                nb, ntau = p.nbands, int(p.qmc()["Ntau"])
                g = []
                for d in dirs:
                    raw = np.fromfile(os.path.join(d, "results.bin"), dtype="<f8")
                    g.append(raw[1:1 + nb * 2 * ntau].reshape((nb, 2, ntau), order="F"))
                moves = nproc * (nwarm + nmeas * ncorr)
                return dict(value=moves / wall, unit="moves/s", cores=nproc, kind="reference",
                            sample="%d processes x (%d warmup + %d x %d measured moves), %s" % (
                                nproc, nwarm, nmeas, ncorr, p.label),
                            gtau_inv_err2_per_s=gtau_inv_err2({"gtau": np.array(g)}) / wall,
                            wall_s=wall)
        except Exception as e:  # reported, then fall back to the port
            print("reference CPU run failed (%s); timing the C oracle instead" % e, file=sys.stderr)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import OracleChain
    o = OracleChain.from_problem(p, seed)
    o.set_taudiffmax(tdm)
    t0 = time.perf_counter()
    o.warmup(nwarm)
    o.measure(nmeas, ncorr)
    wall = time.perf_counter() - t0
    return dict(value=(nwarm + nmeas * ncorr) / wall, unit="moves/s", cores=1, kind="port",
                sample="1 chain x (%d warmup + %d x %d measured moves), %s" % (nwarm, nmeas, ncorr,
                                                                               p.label),
                wall_s=wall)


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


def traffic_from_profiles(workload: str):
    """HBM bytes per sweep-kernel dispatch measured with rocprofv3 --pmc for this
    workload (profiles/sweep_kernel_traffic_<workload>.txt), if recorded."""
    path = os.path.join(ROOT, "profiles", "sweep_kernel_traffic_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_dispatch:"):
                return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--config", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--chains", type=int, default=0)
    ap.add_argument("--ncorr", type=int, default=100)
    ap.add_argument("--seed", type=int, default=43890)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
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
    prob, tdm = make_problem(args.config, args.ncorr)
    _, _, _, nchains_default, capacity = WORKLOADS[args.config]
    nchains = args.chains or nchains_default
    seed = args.seed

    b = run_gpu(prob, tdm, nchains, capacity, args.warmup, args.steps, args.ncorr, seed)
    value = b["moves"] / b["wall"]
    inv_err2_s = b["inv_err2"] / b["wall"]
    kernel_s = b["kernel_ms"] / 1e3
    ach_gbs = b["alg_bytes"] / kernel_s / 1e9
    ach_tfs = b["alg_flops"] / kernel_s / 1e12

    plain = None
    if not args.no_plain:
        psteps = max(32, args.steps // 4)
        p = run_plain(prob, tdm, capacity, args.warmup, psteps, args.ncorr, seed)
        plain = dict(moves_per_s=p["moves"] / p["wall"], gtau_inv_err2_per_s=p["inv_err2"] / p["wall"],
                     sweeps=psteps, wall_s=p["wall"])

    cpu = None
    if not args.no_cpu:
        cprob, _ = make_problem(args.config, args.ncorr)
        cpu = cpu_baseline(cprob, tdm, seed, nwarm=100000, nmeas=10000, ncorr=args.ncorr)

    line = {
        "metric": "MC moves/s (batched chains, 1 GPU)",
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
        "data": "synthetic (semicircular-bath impurity model)",
        "config": {"workload": "%s: %s, beta=%g, %d batched chains, NCorr=%d, TaudiffMax=%g" % (
            args.config, prob.label, prob.beta, nchains, args.ncorr, tdm),
            "chains": nchains, "ncorr": args.ncorr, "mean_order_ops": b["order"]},
        "gtau_inv_err2_per_s": inv_err2_s,
        "plain_single_chain": plain,
        "batched_vs_plain_inv_err2_ratio": (inv_err2_s / plain["gtau_inv_err2_per_s"]) if plain else None,
        "batched_vs_plain_moves_ratio": (value / plain["moves_per_s"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": ach_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_gbs / HBM_PEAK_GBS, "traffic": traffic_from_profiles(args.config),
                     "kernel": "sweep_kernel", "launches": b["launches"],
                     "avg_launch_ms": b["kernel_ms"] / max(1, b["launches"]),
                     "alg_bytes_per_launch": b["alg_bytes"] / max(1, b["launches"]),
                     "fp64_tflops": ach_tfs, "fp64_frac": ach_tfs / FP64_PEAK_TFS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
