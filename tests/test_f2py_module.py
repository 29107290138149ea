"""The Python extension CTQMC (w2dynamics_amd/CTQMC*.so): f2py over the Fortran
host module, built the way src/ctqmc_fortran/CMakeLists.txt builds the reference's
CTQMC from CTQMC.F90.  cthyb_solver.py imports it as
``from w2dyn.auxiliaries.CTQMC import mctqmc as ctqmc`` (cthyb_solver.py:21) and
calls init_paras -> init_rnd -> init_solver -> set_taudiffmax -> ctqmc_warmup ->
ctqmc_measure, then reads the module arrays in get_resultset
(cthyb_solver.py:947-1030).  These tests make the same calls and reads."""

import numpy as np
import pytest

from conftest import load_golden, qmc_of

ENTRY_POINTS = ["init_paras", "init_rnd", "init_solver", "set_taudiffmax", "init_counters", "ctqmc_calibrate",
                "ctqmc_warmup", "ctqmc_measure", "get_mc_config_scalars", "get_mc_config", "set_mc_config",
                "set_empty_mc_config", "dest_ctqmc", "dest_paras"]

# arrays get_resultset reads for every Z-space run (cthyb_solver.py:949-983)
RESULTSET = {"gtau": "gtau", "gtau-full": "gtau_full", "rhist": "rhisto", "lhist": "lhisto", "hist": "histo",
             "ssts-states": "statessuperstates", "energies-eigenstates": "eigenstatesenergies",
             "occbasis-mapping": "occbasismapping", "hist-sst": "outersuperstatehisto",
             "hist-state": "outerstatehisto", "sign-sst": "signhistosuperstates", "sign-state": "signhistostates",
             "contrib-sst": "tracecontribsuperstates", "contrib-state": "tracecontribstates", "occ": "occ",
             "gleg": "gleg", "gleg-full": "gleg_full", "accept-ins": "accadd", "accept-ins4": "accadd4",
             "accept-rem": "accrem", "accept-rem4": "accrem4", "accept-pair-tau": "accpair",
             "accept-glob": "accglob", "accept-shift": "accshift", "accept-flavourchange": "accflavc",
             "sign": "mean_sign", "time-warmup": "time_warmup", "time-simulation": "time_sim",
             "time-sampling": "time_sim_steps"}
# and with MeasDensityMatrix (cthyb_solver.py:1001-1004)
DENSITY = {"densitymatrix": "densitymatrix", "rho1": "rho1", "rho2": "rho2"}


def _module():
    from w2dynamics_amd import CTQMC
    return CTQMC.mctqmc


def test_f2py_module_exposes_the_reference_interface():
    """The extension imports (it loads the in-tree Fortran host and HIP
    libraries) and carries the reference module's entry points and every array
    name get_resultset reads."""
    m = _module()
    for name in ENTRY_POINTS:
        assert callable(getattr(m, name)), name
    for name in list(RESULTSET.values()) + list(DENSITY.values()):
        assert hasattr(m, name), name
    # the C-ABI bindings and internal helpers are not part of its interface
    assert not hasattr(m, "w2d_ctqmc_create") and not hasattr(m, "collect_results")
    # kinds map to double precision, as the reference's .f2py_f2cmap makes them
    assert "array('d')" in m.init_solver.__doc__


# one run through the extension, in a child process (as the Fortran host tests
# run their driver), so that a failure of the module cannot take the test
# process down: argv = fixture name, output .npz
DRIVE = r"""
import sys
import numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(tests)r)
from conftest import load_golden, qmc_of
from w2dynamics_amd import CTQMC
m = CTQMC.mctqmc
z = load_golden(sys.argv[1])
nb = int(z["nbands"])
n2 = 2 * nb
ftau = np.asfortranarray(z["ftau"], dtype=np.float64)
params = str(z["param_string"])
giw_mode = len(sys.argv) > 3 and sys.argv[3] == "giw"
if giw_mode:
    # MeasGiw with 64 frequencies (result["giw-meas"] = ctqmc.giw, cthyb_solver.py:989-990)
    lines = params.split("\n")
    lines = ["MeasGiw = 1 #" if l == "MeasGiw = 0 #" else "Niw = 64 #" if l.startswith("Niw =") else l
             for l in lines]
    params = "\n".join(lines)
m.init_paras(params)
m.init_rnd(int(z["seed"]))
m.init_solver(np.asfortranarray(np.asarray(z["umatrix"]).reshape(n2, n2, n2, n2)), ftau,
              np.asfortranarray(z["muimp"], dtype=np.float64), np.zeros_like(ftau), 1)
assert int(m.aborted) == 0, "init_solver"
m.set_taudiffmax(float(z["taudiff_max_param"]))
m.ctqmc_warmup(int(z["nwarm"]))
m.ctqmc_measure(1, 1)
assert int(m.aborted) == 0, "ctqmc_measure"
out = {"mean_sign": np.array(float(m.mean_sign)), "gtau": m.gtau, "gleg": m.gleg, "occ": m.occ,
       "densitymatrix": m.densitymatrix, "histo": m.histo,
       "acc": np.array([m.accadd[0], m.accrem[0], float(m.accglob), float(m.accshift)]),
       "outer_sst_histo": m.outersuperstatehisto}
# get_resultset (cthyb_solver.py:947-1030): its dictionary of copies
keys = dict(%(resultset)r)
if int(qmc_of(z).get("MeasDensityMatrix", "0")) != 0:
    keys.update(%(density)r)
result = dict((k, np.copy(getattr(m, v))) for k, v in keys.items())
norb = result["gtau"].shape[0]
if "rho2" in result:
    result["rho2"] = result["rho2"].reshape((norb, 2) * 4)
    result["rho1"] = result["rho1"].reshape((norb, 2) * 2)
if giw_mode:
    result["giw-meas"] = np.copy(m.giw)
for k, v in result.items():
    out["rs_" + k] = np.asarray(v)
if len(sys.argv) > 3 and sys.argv[3] == "reuse":
    # ReuseMCConfig (cthyb_solver.py:24-86, 328-332): the last configuration
    # goes into the next solve right after init_solver
    n, osst, ostate = m.get_mc_config_scalars()
    cfg = m.get_mc_config(n) if n > 0 else None
    m.dest_ctqmc()
    m.init_solver(np.asfortranarray(np.asarray(z["umatrix"]).reshape(n2, n2, n2, n2)), ftau,
                  np.asfortranarray(z["muimp"], dtype=np.float64), np.zeros_like(ftau), 2)
    assert int(m.aborted) == 0, "init_solver 2"
    if cfg is not None:
        m.set_mc_config(n, *cfg[:5], cfg[5], cfg[6])
    else:
        m.set_empty_mc_config(osst, ostate)
    assert int(m.aborted) == 0, "set_mc_config"
    n2_, osst2, _ = m.get_mc_config_scalars()
    back = m.get_mc_config(n2_) if n2_ > 0 else None
    out["reuse_n"] = np.array([n, n2_, osst, osst2])
    if cfg is not None:
        for i, k in enumerate(("taus", "orbs", "spins", "cas", "hashybs")):
            out["reuse_set_" + k] = np.asarray(cfg[i])
            out["reuse_back_" + k] = np.asarray(back[i])
    m.set_taudiffmax(float(z["taudiff_max_param"]))
    m.ctqmc_warmup(100)
    m.ctqmc_measure(1, 1)
    assert int(m.aborted) == 0, "second solve"
    out["reuse_gtau"] = np.asarray(m.gtau)
m.dest_ctqmc()
m.dest_paras()
np.savez(sys.argv[2], **out)
"""


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_run", "c2_run"])
def test_f2py_module_reproduces_reference_run(tmp_path, name):
    """One chain driven through the extension reproduces the reference's own
    run of the golden fixture (the same Markov chain move for move): the
    observables agree to rounding, and get_resultset's dictionary builds."""
    import os
    import subprocess
    import sys
    from conftest import ROOT
    z = load_golden(name)
    print("f2py module run", name, flush=True)
    script = DRIVE % dict(root=ROOT, tests=os.path.join(ROOT, "tests"), resultset=RESULTSET, density=DENSITY)
    out = str(tmp_path / "out.npz")
    r = subprocess.run([sys.executable, "-c", script, name, out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.load(out)
    for k in ("mean_sign", "gtau", "gleg", "occ", "densitymatrix", "histo", "acc", "outer_sst_histo"):
        a, b = np.asarray(got[k]), np.asarray(z["res_" + k])
        assert a.shape == b.shape, (k, a.shape, b.shape)
        assert np.allclose(a, b, rtol=1e-8, atol=1e-10, equal_nan=True), (k, np.abs(a - b).max())
    q = qmc_of(z)
    norb = int(z["nbands"])
    assert got["rs_gtau"].shape == (norb, 2, int(q["Ntau"]))
    assert got["rs_hist-sst"].shape == (int(z["model_nsst"]),)
    for k in RESULTSET:
        assert "rs_" + k in got.files, k


@pytest.mark.gpu
def test_f2py_module_reuses_the_mc_configuration(tmp_path):
    """ReuseMCConfig through the extension: the configuration read after one
    solve (get_mc_config_scalars, get_mc_config) is set right after the next
    init_solver (set_mc_config(N, ...)), reads back unchanged, and the second
    solve runs from it."""
    import os
    import subprocess
    import sys
    from conftest import ROOT
    script = DRIVE % dict(root=ROOT, tests=os.path.join(ROOT, "tests"), resultset=RESULTSET, density=DENSITY)
    out = str(tmp_path / "out.npz")
    print("f2py module reuse", flush=True)
    r = subprocess.run([sys.executable, "-c", script, "c2_run", out, "reuse"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.load(out)
    n, n_back, osst, osst_back = got["reuse_n"]
    assert n > 0 and n_back == n and osst_back == osst
    for k in ("taus", "orbs", "spins", "cas", "hashybs"):
        assert np.array_equal(got["reuse_set_" + k], got["reuse_back_" + k]), k
    assert np.all(got["reuse_set_hashybs"] == -1)
    assert np.all(np.isfinite(got["reuse_gtau"]))


@pytest.mark.gpu
def test_f2py_module_measures_giw(tmp_path):
    """MeasGiw through the extension: ``giw`` (NBands, 2, 2 Niw) of the golden
    run equals the oracle's G(i w_n) of the same Markov chain."""
    import os
    import subprocess
    import sys
    from conftest import ROOT
    from oracle import OracleChain
    z = load_golden("c2_run")
    script = DRIVE % dict(root=ROOT, tests=os.path.join(ROOT, "tests"), resultset=RESULTSET, density=DENSITY)
    out = str(tmp_path / "out.npz")
    print("f2py module giw", flush=True)
    r = subprocess.run([sys.executable, "-c", script, "c2_run", out, "giw"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.load(out)["rs_giw-meas"]
    q = qmc_of(z)
    q["MeasGiw"], q["Niw"] = "1", "64"
    ch = OracleChain(z["umatrix"], z["ftau"], z["muimp"], float(z["beta"]), q, int(z["seed"]))
    ch.set_taudiffmax(float(q["TaudiffMax"]))
    ch.warmup(int(z["nwarm"]))
    ch.measure(int(q["Nmeas"]), int(q["NCorr"]))
    want = ch.results()["giw"]
    assert got.shape == want.shape == (int(z["nbands"]), 2, 128)
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()


def test_f2py_module_reports_library_errors(tmp_path):
    """An unsupported parameter (state sampling): init_solver reports the
    library's status and sets aborted = 2 instead of stopping the process, as
    the Fortran host does; the module stays usable (dest_ctqmc, dest_paras)."""
    import os
    import subprocess
    import sys
    from conftest import ROOT
    script = r"""
import sys
import numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(tests)r)
from conftest import load_golden
from w2dynamics_amd import CTQMC
m = CTQMC.mctqmc
z = load_golden("c1_run")
lines = [("statesampling = 1" if ln.split("=")[0].strip() == "statesampling" else ln)
         for ln in str(z["param_string"]).split("\n")]
assert "statesampling = 1" in lines
m.init_paras("\n".join(lines))
m.init_rnd(1)
nb = int(z["nbands"]); n2 = 2 * nb
ftau = np.asfortranarray(z["ftau"], dtype=np.float64)
m.init_solver(np.asfortranarray(np.asarray(z["umatrix"]).reshape(n2, n2, n2, n2)), ftau,
              np.asfortranarray(z["muimp"], dtype=np.float64), np.zeros_like(ftau), 1)
print("aborted", int(m.aborted))
m.dest_ctqmc()
m.dest_paras()
""" % dict(root=ROOT, tests=os.path.join(ROOT, "tests"))
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "aborted 2" in r.stdout
    assert "w2d_ctqmc_create returned status" in r.stderr
