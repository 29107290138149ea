"""Fortran host module MCTQMC (w2dynamics_amd/fortran/ctqmc_host.F90).

The test driver tests/fortran/host_run.F90 calls the module exactly as the
reference driver oracle/ref_driver.F90 calls the reference module (init_paras ->
init_rnd -> init_solver -> set_taudiffmax -> ctqmc_warmup -> ctqmc_measure(1,1),
cthyb_solver.py:240-342) and writes results.bin in the same layout.  On the GPU,
its output for the golden run fixtures (produced by the reference) must agree with
the reference's own results: the single chain reproduces the reference's Markov
chain move for move, so the observables agree to rounding.
"""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

DRIVER = os.path.join(ROOT, "tests", "fortran", "host_run")
FLIB = os.path.join(ROOT, "w2dynamics_amd", "libw2dyn_ctqmc_fortran.so")

ENTRY_POINTS = ["init_paras", "init_rnd", "init_solver", "set_taudiffmax", "ctqmc_warmup",
                "ctqmc_measure", "init_counters", "get_mc_config_scalars", "get_mc_config",
                "set_mc_config", "set_empty_mc_config", "dest_ctqmc", "dest_paras"]


def write_inputs_from_fixture(d, z, nchains=1):
    """Input directory of the reference driver for a golden run fixture."""
    nb = int(z["nbands"])
    ftau = np.asarray(z["ftau"])
    with open(os.path.join(d, "run.txt"), "w") as f:
        f.write("run\n%d\n%d\n%d\n%d\n0\n%.17g\n%d\n" % (nb, ftau.shape[-1], int(z["seed"]),
                                                       int(z["nwarm"]),
                                                       float(z["taudiff_max_param"]), nchains))
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write(str(z["param_string"]))
    n = 2 * nb
    np.asarray(z["umatrix"]).reshape(n, n, n, n).ravel(order="F").tofile(os.path.join(d, "umatrix.bin"))
    np.asfortranarray(ftau).ravel(order="F").tofile(os.path.join(d, "ftau.bin"))
    np.asfortranarray(z["muimp"]).ravel(order="F").tofile(os.path.join(d, "muimp.bin"))


def read_results(d, z):
    nb, ntau, nleg = int(z["nbands"]), int(z["ntau"]), int(z["nlegmax"])
    nst = 4 ** nb
    maxh = np.asarray(z["res_histo"]).shape[-1] - 1
    nsst = int(z["model_nsst"])
    raw = np.fromfile(os.path.join(d, "results.bin"), dtype="<f8")
    out, off = {}, 0
    for key, n, shape in (("mean_sign", 1, ()), ("gtau", nb * 2 * ntau, (nb, 2, ntau)),
                          ("gleg", nb * 2 * nleg, (nb, 2, nleg)), ("occ", (2 * nb) ** 2, (nb, 2, nb, 2)),
                          ("densitymatrix", nst * nst, (nst, nst)),
                          ("histo", nb * 2 * (maxh + 1), (nb, 2, maxh + 1)), ("acc", 4, (4,)),
                          ("outer_sst_histo", nsst, (nsst,))):
        a = raw[off:off + n]
        off += n
        out[key] = a.reshape(shape, order="F") if shape else a[0]
    assert off == raw.size
    return out


def test_fortran_host_built_and_exports_entry_points():
    assert os.path.exists(DRIVER) and os.path.exists(FLIB), "run __graft_entry__.build() first"
    syms = subprocess.run(["nm", "-D", "--defined-only", FLIB], capture_output=True, text=True,
                          check=True).stdout.lower()
    for name in ENTRY_POINTS:
        assert "mmctqmcp" + name in syms, name     # flang module-procedure mangling


def test_fortran_host_reports_library_errors(tmp_path):
    """Without a usable GPU, or with an unsupported parameter, init_solver reports
    the library's status and the driver stops with its own code instead of
    crashing."""
    z = load_golden("c1_run")
    write_inputs_from_fixture(str(tmp_path), z)
    txt = (tmp_path / "params.txt").read_text().split("\n")
    txt = [("statesampling = 1" if ln.split("=")[0].strip() == "statesampling" else ln) for ln in txt]
    assert "statesampling = 1" in txt
    (tmp_path / "params.txt").write_text("\n".join(txt))   # unsupported by the library: status 1
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "w2d_ctqmc_create returned status" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_run", "c2_run"])
def test_fortran_host_reproduces_reference_run(tmp_path, name):
    z = load_golden(name)
    write_inputs_from_fixture(str(tmp_path), z)
    print("fortran host run", name, flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    res = read_results(str(tmp_path), z)
    for k in ("mean_sign", "gtau", "gleg", "occ", "densitymatrix", "histo", "acc", "outer_sst_histo"):
        a, b = np.asarray(res[k]), np.asarray(z["res_" + k])
        assert a.shape == b.shape, k
        assert np.allclose(a, b, rtol=1e-8, atol=1e-10, equal_nan=True), (k, np.abs(a - b).max())


@pytest.mark.gpu
def test_fortran_host_batched_chains_pooled(tmp_path):
    """n_chains > 1: module arrays are the normalisation of the sums over all
    chains (the sign-ratio estimate), identical to the Python host's pooled
    estimate over the same chains."""
    from w2dynamics_amd.ctqmc import CtqmcGpu
    z = load_golden("c2_run")
    write_inputs_from_fixture(str(tmp_path), z, nchains=8)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    res = read_results(str(tmp_path), z)
    q = str(z["param_string"])
    gpu = CtqmcGpu(q, z["umatrix"], z["ftau"], z["muimp"], nchains=8, seed=int(z["seed"]))
    gpu.set_taudiffmax(float(z["taudiff_max_param"]))
    gpu.warmup(int(z["nwarm"]))
    gpu.measure(int(z["nmeas"]))
    py = gpu.aggregate(ngroups=8)
    gpu.close()
    for k in ("gtau", "gleg", "occ", "histo"):
        assert np.allclose(res[k], py[k], rtol=1e-10, atol=1e-12), k
    assert np.isclose(res["mean_sign"], py["mean_sign"], rtol=1e-12)
