"""The allocation-free group sums (w2d_ctqmc_get_group_sums_noalloc) against
the allocating entry point, through the C-ABI, on two calls in a row (the
second with a smaller handle and fewer groups), and the rejection of a request
larger than the static scratch."""

import ctypes as C

import numpy as np
import pytest

from w2dynamics_amd import problems as P
from w2dynamics_amd.ctqmc import CtqmcGpu, _p

pytestmark = pytest.mark.gpu


def _sums(gpu, ng, noalloc):
    nf = 2 * gpu.nbands
    m = gpu.model()
    dmsize = int((m["sizes"].astype(np.int64) ** 2).sum())
    out = dict(gtau=np.full((ng, nf, gpu.ntau), np.nan), gleg=np.full((ng, nf, gpu.nleg), np.nan),
               dm=np.full((ng, dmsize), np.nan), histo=np.full((ng, nf, gpu.maxhisto + 1), np.nan),
               outer=np.full((ng, gpu.nsst), np.nan), scalars=np.full((ng, 14), np.nan),
               nused=np.full(ng, -1, np.int32))
    fn = gpu.lib.w2d_ctqmc_get_group_sums_noalloc if noalloc else gpu.lib.w2d_ctqmc_get_group_sums
    fn.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 7
    st = fn(gpu.h, ng, *[_p(out[k]) for k in ("gtau", "gleg", "dm", "histo", "outer", "scalars", "nused")])
    return st, out


def _run(prob, nch, seed, nmeas):
    gpu = CtqmcGpu.from_problem(prob, nchains=nch, seed=seed)
    gpu.set_taudiffmax(3.0)
    gpu.warmup(1000)
    gpu.measure(nmeas)
    return gpu


def test_noalloc_group_sums_match_allocating_entry_point():
    prob = P.kanamori_2orb(beta=10.0, TaudiffMax=3.0, Nmeas=50, NCorr=20, Ntau=100, NLegMax=20)
    for nch, ng, seed in ((96, 12, 3), (20, 5, 4)):      # second call: smaller handle, fewer groups
        print("noalloc group sums", nch, "chains", ng, "groups", flush=True)
        gpu = _run(prob, nch, seed, 50)
        st_a, a = _sums(gpu, ng, noalloc=False)
        st_b, b = _sums(gpu, ng, noalloc=True)
        assert st_a == 0 and st_b == 0
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        assert int(b["nused"].sum()) == nch
        assert np.isclose(b["scalars"][:, 1].sum(), 50 * nch)
        gpu.close()


def test_noalloc_group_sums_reject_requests_beyond_the_scratch():
    # 4096-bin G(tau) per flavour: about 25 K doubles per chain, so 160 groups need > 16 MiB
    prob = P.kanamori_2orb(beta=10.0, TaudiffMax=3.0, Nmeas=4, NCorr=20, Ntau=4096, NLegMax=20)
    gpu = _run(prob, 256, 5, 4)
    st, _ = _sums(gpu, 160, noalloc=True)
    assert st == 3                                        # W2D_ERR_CAPACITY
    st_a, a = _sums(gpu, 160, noalloc=False)              # the allocating one still serves it
    assert st_a == 0 and int(a["nused"].sum()) == 256
    gpu.close()
