"""The reference's own Fortran host driving the GPU step (oracle/_ref/dgswem_gpu).

oracle/build_dropin.sh links every object of the reference's serial solver except DG_timestep.o
with dgswem_amd/fortran/dg_timestep_gpu.F90 (DG_TIMESTEP calling the C ABI through
dgswem_amd/fortran/gpu_interface.F90) and libdgswem_hip.so.  The binary reads the same
fort.14/15/dg/13/22 and writes the same fort.63, maxele.63 and DG hot start as dgswem_serial.
The nodal values and ETAMAX stay on the GPU and are copied back only on output steps; inputs the
library does not implement stop the run with a message instead of running without their terms.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden, load_case
from dgswem_amd import fortio, mesh

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "oracle", "_ref", "dgswem_gpu")


def run(d):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/dgswem_gpu is not built (needs the reference sources)")
    r = subprocess.run([BIN], cwd=d, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "CPU_TIME" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]


@pytest.mark.parametrize("name", ["quarter_annular", "wetdry", "mass_conservation"])
def test_dropin_reproduces_reference_fort63(name, tmp_path):
    d = str(tmp_path)
    for f in ("fort.14", "fort.15", "fort.dg"):
        shutil.copy(os.path.join(GOLDEN, name, f), d)
    if name == "wetdry":
        case = fortio.barriers_as_land(fortio.read_case(d))
        fortio.write_fort14(os.path.join(d, "fort.14"), case, "wetdry-land")
    print(f"dropin {name}", flush=True)
    run(d)
    _, _, got = fortio.read_fort63(os.path.join(d, "fort.63"))
    ref = golden(f"{name}_fort63_last")["eta"]
    np.testing.assert_array_equal(got[-1] == -99999.0, ref == -99999.0)
    np.testing.assert_allclose(got[-1], ref, rtol=0, atol=2e-8 * np.abs(ref).max())


def test_dropin_hot_start_matches_reference_fp64(tmp_path):
    """Config-1 basin, 1000 steps: the DG hot start written by the drop-in vs dgswem_serial's."""
    d = str(tmp_path)
    case = mesh.structured_basin(nsteps=1000)
    mesh.write_run_dir(d, case, dghot=1, dghotspool=1000, nspoolge=1000)
    run(d)
    hs = fortio.read_dg_hotstart(d, 3, case.ne)
    ref = golden("c1_step1000")
    np.testing.assert_array_equal(hs["wdflg"], ref["wdflg"])
    for k in ("ze", "qx", "qy"):
        rel = np.linalg.norm(hs[k] - ref[k]) / np.linalg.norm(ref[k])
        assert rel <= 1e-10, (k, rel)


def test_dropin_maxele_matches_reference(tmp_path):
    """maxele.63 after 100 steps of the wetting/drying Delaunay basin, against dgswem_serial's
    (printed f20.8).  ETAMAX is raised on the GPU after every wet element's contribution to ETA2,
    visiting each node's elements in the host's ELETAB order: with another order the partial sums
    differ and some nodes' maxima with them (the oracle, which visits elements in ascending order,
    differs from this file at 7 nodes)."""
    d = str(tmp_path)
    case = mesh.delaunay_basin(n_elem=4000, seed=1, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 30.0),
                               cfl=0.3)
    case.params.nt = 100
    mesh.write_run_dir(d, case, dghot=0, dghotspool=100, nspoolge=100)
    print("dropin maxele delaunay4k", flush=True)
    run(d)
    got = np.loadtxt(os.path.join(d, "maxele.63"), skiprows=1)
    ref = golden("delaunay4k_maxele_step100")
    np.testing.assert_array_equal(got, ref["etamax"])
    _, it, eta = fortio.read_fort63(os.path.join(d, "fort.63"))
    assert it[-1] == 100
    np.testing.assert_array_equal(eta[-1] == -99999.0, ref["eta"] == -99999.0)
    np.testing.assert_allclose(eta[-1], ref["eta"], rtol=0, atol=2e-8 * np.abs(ref["eta"]).max())


@pytest.mark.parametrize("name", ["storm3k", "tipot2k", "tipotsal2k", "c1eseries", "c1fseries"])
def test_dropin_forcing_matches_reference_fp64(tmp_path, name):
    """Forcing the drop-in hands to the library from the host's modules: NWS=8 (it reads fort.22
    itself, FORMAT 14), the tidal potential (NTIP=1, and NTIP=2 with fort.24's self-attraction
    terms), a fort.19 elevation series (NBFR=0) and a fort.20 flux series (NFFR=0; the first two
    records come from the host's QNIN1/QNIN2, the rest from unit 20).  The DG hot start at the end
    matches dgswem_serial's."""
    d = str(tmp_path)
    for f in ("fort.14", "fort.15", "fort.13", "fort.dg", "fort.22", "fort.24", "fort.19", "fort.20"):
        if os.path.exists(os.path.join(GOLDEN, name, f)):
            shutil.copy(os.path.join(GOLDEN, name, f), d)
    case = fortio.read_case(d)
    steps = case.params.nt
    assert case.dg.dghot == 1 and case.dg.dghotspool == steps
    print(f"dropin {name} {steps} steps", flush=True)
    run(d)
    hs = fortio.read_dg_hotstart(d, case.dofh, case.ne)
    ref = golden(f"{name}_step{steps}")
    np.testing.assert_array_equal(hs["wdflg"], ref["wdflg"])
    for k in ("ze", "qx", "qy"):
        rel = np.linalg.norm(hs[k] - ref[k]) / np.linalg.norm(ref[k])
        assert rel <= 1e-10, (k, rel)
