"""Launch machinery on the GPU: every admissible trace engine reproduces the
oracle's move sequence; graph-captured and ordered launches change no chain's
results; the autotuner times the engines and keeps a valid one."""

import numpy as np
import pytest

from oracle import OracleChain
from test_gpu_parity import _compare
from w2dynamics_amd import problems as P
from w2dynamics_amd.ctqmc import CtqmcGpu

pytestmark = pytest.mark.gpu

ENGINES = {
    "c1_nm4": (lambda: P.anderson_1orb(TaudiffMax=2.0), 4, 4000),
    "c1_nm16": (lambda: P.anderson_1orb(TaudiffMax=2.0), 16, 4000),
    "c1_nm128": (lambda: P.anderson_1orb(TaudiffMax=2.0), 128, 3000),
    "c2_nm16": (lambda: P.kanamori_2orb(beta=20.0, TaudiffMax=4.0), 16, 4000),
    "c2_nm128": (lambda: P.kanamori_2orb(beta=20.0, TaudiffMax=4.0), 128, 3000),
    "c3_nm128": (lambda: P.kanamori_3orb(beta=20.0, TaudiffMax=5.0), 128, 2000),
}

# The MFMA engine on models whose blocks fit 16 states: a measurement pass
# with it (3-orbital model, after the autotuner had switched to it) ended in a
# GPU memory fault whose cause is not found; the library refuses the
# combination, and these cases stay skipped until the cause is shown.
MFMA_SMALL_BLOCKS = ("MFMA engine on models with blocks of at most 16 states is refused by the "
                     "library: a measurement pass with it faulted the GPU, cause not found")


def _engine_cases():
    return [pytest.param(k, marks=pytest.mark.skip(reason=MFMA_SMALL_BLOCKS)) if ENGINES[k][1] == 128 else k
            for k in ENGINES]


@pytest.mark.parametrize("case", _engine_cases())
def test_every_engine_matches_oracle(case):
    mk, nm, n = ENGINES[case]
    prob = mk()
    tdm = float(prob.qmc()["TaudiffMax"])
    seed = 515
    print("engine case", case, flush=True)
    gpu = CtqmcGpu.from_problem(prob, nchains=2, seed=seed)
    gpu.set_engine(nm)
    assert gpu.engine()[0] == nm
    gpu.set_taudiffmax(tdm)
    rec = gpu.moves_log(0, n)
    ora = OracleChain.from_problem(prob, seed)
    ora.set_taudiffmax(tdm)
    _compare(rec, ora.moves(n), case)
    gpu.close()


def _run(prob, tdm, **launch):
    gpu = CtqmcGpu.from_problem(prob, nchains=96, seed=4242)
    gpu.set_launch(**launch)
    gpu.set_taudiffmax(tdm)
    gpu.warmup(3000)
    gpu.measure(60)
    acc = gpu.accumulators()
    cfg = [gpu.get_config(c) for c in (0, 50, 95)]
    gpu.close()
    return acc, cfg


# Chain ordering (order_chains) is off by default: every GPU memory fault seen
# with this library came from runs with it on (trace engine for blocks above 16
# states); its variants stay skipped until that is explained.
ORDERING_OFF = "longest-first chain ordering is switched off: runs with it on ended in GPU memory faults"
LAUNCH_OPTS = [
    pytest.param(dict(graphs=True, order_chains=False), id="graphs"),
    pytest.param(dict(graphs=False, order_chains=True), id="order", marks=pytest.mark.skip(reason=ORDERING_OFF)),
    pytest.param(dict(graphs=True, order_chains=True), id="graphs_order",
                 marks=pytest.mark.skip(reason=ORDERING_OFF)),
]


@pytest.mark.parametrize("opts", LAUNCH_OPTS)
def test_graphs_and_chain_order_change_no_result(opts):
    """Many short launches (2 sweeps each) with graph capture (and chain
    ordering): identical configurations, accumulators to rounding."""
    prob = P.kanamori_2orb(beta=20.0, TaudiffMax=4.0, NCorr=50)
    ref, rcfg = _run(prob, 4.0, graphs=False, order_chains=False, moves_per_launch=100)
    print("launch options", opts, flush=True)
    acc, cfg = _run(prob, 4.0, moves_per_launch=100, **opts)
    for a, b in zip(cfg, rcfg):
        assert a["n"] == b["n"] and np.array_equal(a["taus"], b["taus"]), opts
    for k in ("gtau", "gleg", "dm", "histo", "scalars"):
        assert np.allclose(acc[k], ref[k], rtol=1e-12, atol=1e-12), (opts, k)


@pytest.mark.skip(reason=MFMA_SMALL_BLOCKS)
def test_autotuner_picks_an_admissible_engine():
    prob = P.kanamori_3orb(beta=20.0, TaudiffMax=5.0)
    gpu = CtqmcGpu.from_problem(prob, nchains=256, seed=99)
    gpu.set_taudiffmax(5.0)
    print("autotune c3", flush=True)
    rates = gpu.autotune(400)
    print(rates, flush=True)
    assert set(k for k in rates if k != "chosen") == {16, 128}
    # 256 chains x 400 moves on one GPU: rates of order 1e5-1e8 moves/s, not
    # the rate of an empty pass
    assert all(1e3 < rates[k] < 1e9 for k in (16, 128)), rates
    assert rates["chosen"] == max((k for k in rates if k != "chosen"), key=lambda k: rates[k])
    assert gpu.engine()[0] == rates["chosen"]
    gpu.measure(20)
    assert np.all(gpu.chain_status()["status"] == 0)
    gpu.close()


def test_autotuner_times_the_small_block_engines():
    """2-orbital model (blocks of up to 4 states): engines 4 and 16 timed,
    the MFMA engine not admissible, the faster one kept; a measurement with
    the chosen engine leaves every chain valid."""
    prob = P.kanamori_2orb(beta=20.0, TaudiffMax=4.0)
    gpu = CtqmcGpu.from_problem(prob, nchains=256, seed=99)
    gpu.set_taudiffmax(4.0)
    print("autotune c2", flush=True)
    rates = gpu.autotune(400)
    print(rates, flush=True)
    assert set(k for k in rates if k != "chosen") == {4, 16}
    assert all(1e3 < rates[k] < 1e9 for k in (4, 16)), rates
    assert rates["chosen"] == max((k for k in rates if k != "chosen"), key=lambda k: rates[k])
    assert gpu.engine()[0] == rates["chosen"]
    gpu.warmup(200)
    gpu.measure(20)
    assert np.all(gpu.chain_status()["status"] == 0)
    gpu.close()


def test_mfma_engine_refused_for_small_blocks():
    """set_engine(128) on a model whose blocks fit 16 states is an argument
    error and leaves the engine unchanged."""
    from w2dynamics_amd.ctqmc import CtqmcError
    gpu = CtqmcGpu.from_problem(P.kanamori_3orb(beta=20.0, TaudiffMax=5.0), nchains=4, seed=1)
    nm0 = gpu.engine()[0]
    with pytest.raises(CtqmcError):
        gpu.set_engine(128)
    assert gpu.engine()[0] == nm0
    gpu.close()
