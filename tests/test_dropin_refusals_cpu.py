"""The Fortran drop-in (oracle/_ref/dgswem_gpu) refuses inputs whose terms the GPU step does not
have, before any GPU work: READ_INPUT and PREP_DG run on the host, then the first DG_TIMESTEP
hands the parameters to dg_gpu_init, which checks them before it touches a device and fails; the
drop-in STOPs with the library's message instead of running without those terms."""
import os
import subprocess

import pytest

from conftest import ROOT
from dgswem_amd import mesh

BIN = os.path.join(ROOT, "oracle", "_ref", "dgswem_gpu")


def _set_dg_key(d, key, value):
    p = os.path.join(d, "fort.dg")
    lines = open(p).read().split("\n")
    for i, l in enumerate(lines):
        if l.split("!")[0].split("=")[0].strip() == key:
            lines[i] = f"{key} = {value}"
    open(p, "w").write("\n".join(lines))


@pytest.mark.parametrize("key,value,message", [
    ("artdif", "1", "artdif=1: LDG artificial diffusion is not supported"),
    ("padapt", "1", "padapt=1: p-adaptivity is not supported"),
    ("pl", "0", "fort.dg pl=0 differs from ph=1"),
])
def test_dropin_refuses_what_it_does_not_run(tmp_path, key, value, message):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/dgswem_gpu is not built (needs the reference sources)")
    d = str(tmp_path)
    mesh.write_run_dir(d, mesh.structured_basin(nsteps=5))
    _set_dg_key(d, key, value)
    r = subprocess.run([BIN], cwd=d, capture_output=True, text=True, timeout=120)
    assert "DG GPU ERROR in dg_gpu_init" in r.stdout and message in r.stdout, r.stdout[-2000:]
    assert "CPU_TIME" not in r.stdout and "TIME STEP =       1" not in r.stdout


@pytest.mark.parametrize("key,value,message", [
    ("pl", "1", "fort.dg pl=1 differs from ph=2"),
    ("padapt", "1", "padapt=1"),
    ("artdif", "1", "artdif=1"),
])
def test_read_case_refuses_fort_dg_options_it_does_not_run(tmp_path, key, value, message):
    """With padapt=0 the reference runs at order pl (src/prep_DG.F:90,104,124): a pl=1, ph=2 fort.dg
    is refused rather than run at p=2."""
    from dgswem_amd import fortio
    d = str(tmp_path)
    mesh.write_run_dir(d, mesh.structured_basin(nsteps=5, p=2))
    _set_dg_key(d, key, value)
    with pytest.raises(NotImplementedError, match=message):
        fortio.read_case(d)


@pytest.mark.parametrize("field,value,message", [
    ("ntip", 3, "NTIP=3"), ("nws", 1, "NWS=1"), ("artdif", 1, "artdif=1"), ("padapt", 1, "padapt=1"),
    ("pl", 2, "pl=2 differs from ph=1"), ("load_quadratic_fric", 1, "quadratic friction"),
])
def test_library_refuses_parameters_before_touching_a_device(field, value, message):
    """dg_gpu_init checks the parameters first, so this runs without a GPU."""
    import ctypes as C
    from dgswem_amd import gpu
    case = mesh.structured_basin(nx=4, ny=2, nsteps=5)
    prm = gpu.params_of(case)
    setattr(prm, field, value)
    L = gpu.load_library()
    ed = gpu.create_edge_data(case)
    h = C.c_int(-1)
    f = lambda a: gpu._f(a)
    i = lambda a: gpu._i(a)
    rc = L.dg_gpu_init(C.byref(prm), case.np_, case.ne, f(case.x), f(case.y), f(case.dp),
                       f(case.sfac), f(case.corif), None, i(case.nm.T.reshape(-1)), ed["nedges"],
                       i(ed["nedno"].reshape(-1)), i(ed["nedel"].reshape(-1)),
                       i(ed["nedsd"].reshape(-1)), len(ed["niedn"]), i(ed["niedn"]),
                       len(ed["nledn"]), i(ed["nledn"]), len(ed["needn"]), i(ed["needn"]),
                       len(ed["nfedn"]), i(ed["nfedn"]), C.byref(h))
    assert rc != 0
    assert message in L.dg_gpu_last_error().decode()
