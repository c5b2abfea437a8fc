"""Generate the FP64 golden fixtures under tests/golden/ by running the reference's serial solver.

Run in the build container (needs oracle/_ref/dgswem_serial, built by oracle/build_ref.sh from
/root/reference).  For each case it writes a run directory, switches on the DG hot start
(dghot=1, dghotspool=N) so dgswem_serial dumps its modal state with 17 significant digits
(src/write_results.F:1199-1245), runs N steps, and stores the state as <case>_stepN.npz.
Cases taken from the reference's own tests also keep their input files and the last snapshot of
the reference's fort.63.true (<case>_fort63_last.npz).

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py c1_8640    # only the named groups
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from dgswem_amd import fortio, mesh  # noqa: E402

REF = os.environ.get("DGSWEM_REFERENCE", "/root/reference")
BIN = os.path.join(ROOT, "oracle", "_ref", "dgswem_serial")

REFERENCE_TESTS = ["quarter_annular", "wetdry", "mass_conservation"]


def run_ref(run_dir: str) -> None:
    r = subprocess.run([BIN], cwd=run_dir, capture_output=True, text=True)
    # a normal run also ends in STOP (src/dgswem.F:443), after the CPU_TIME line
    if r.returncode != 0 or "CPU_TIME" not in r.stdout:
        raise RuntimeError(f"dgswem_serial failed in {run_dir}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")


def set_dghot(run_dir: str, nsteps: int) -> None:
    """Rewrite fort.dg keys dghot/dghotspool (keyword format) and RNDAY in fort.15 for nsteps."""
    p = os.path.join(run_dir, "fort.dg")
    lines = open(p).read().split("\n")
    out = []
    for l in lines:
        k = l.split("!")[0].split("=")[0].strip()
        if k == "dghot":
            l = "dghot = 1"
        elif k == "dghotspool":
            l = f"dghotspool = {nsteps}"
        out.append(l)
    open(p, "w").write("\n".join(out))
    # RNDAY is record 24 of fort.15 when NWP=0 (the reference test cases); locate it by reading
    case = fortio.read_case(run_dir)
    dt = case.params.dtdp
    rnday = nsteps * dt / 86400.0
    f15 = open(os.path.join(run_dir, "fort.15"), newline="").read().split("\n")
    # records: title, runid, NFOVER..NWP (11), [attr names], NCOR NTIP NWS NRAMP G TAU0 DT STATIM REFTIM RNDAY
    idx = 2 + 11 + case.params.nwp + 9
    f15[idx] = repr(rnday) + "   ! RNDAY"
    open(os.path.join(run_dir, "fort.15"), "w", newline="").write("\n".join(f15))
    assert fortio.read_case(run_dir).params.nt == nsteps


def dump(run_dir: str, case, name: str, nsteps: int) -> None:
    hs = fortio.read_dg_hotstart(run_dir, case.dofh, case.ne)
    assert hs["it"] == nsteps, (hs["it"], nsteps)
    np.savez_compressed(os.path.join(HERE, f"{name}_step{nsteps}.npz"), ze=hs["ze"], qx=hs["qx"],
                        qy=hs["qy"], wdflg=hs["wdflg"], it=np.int64(nsteps))
    print(f"  {name}: step {nsteps} written")


def generated_case(name: str, case, steps) -> None:
    for n in steps:
        case.params.nt = n
        with tempfile.TemporaryDirectory() as d:
            mesh.write_run_dir(d, case, dghot=1, dghotspool=n, nspoolge=n)
            if name.startswith(("delaunay", "storm")) and n == steps[0]:
                keep = os.path.join(HERE, name)
                os.makedirs(keep, exist_ok=True)
                for f in ("fort.14", "fort.15", "fort.13", "fort.dg", "fort.22"):
                    if os.path.exists(os.path.join(d, f)):
                        shutil.copy(os.path.join(d, f), keep)
            run_ref(d)
            dump(d, fortio.read_case(d), name, n)


def reference_case(name: str, steps) -> None:
    src = os.path.join(REF, "tests", name)
    keep = os.path.join(HERE, name)
    os.makedirs(keep, exist_ok=True)
    for f in ("fort.14", "fort.15", "fort.dg"):
        shutil.copyfile(os.path.join(src, f), os.path.join(keep, f))
    # the reference's test compares the last fort.63 snapshot only (tests/util.py:30-53)
    t, it, vals = fortio.read_fort63(os.path.join(src, "fort.63.true"))
    np.savez_compressed(os.path.join(HERE, f"{name}_fort63_last.npz"), time=t[-1], it=it[-1],
                        eta=vals[-1])
    for n in steps:
        with tempfile.TemporaryDirectory() as d:
            for f in ("fort.14", "fort.15", "fort.dg"):
                shutil.copy(os.path.join(src, f), d)
            case = fortio.read_case(d)
            if name == "wetdry":
                # barrier segments become land segments; below the barrier height (never reached in
                # this case) a barrier edge reflects like a land edge.
                case = fortio.barriers_as_land(case)
                fortio.write_fort14(os.path.join(d, "fort.14"), case, "wetdry-land")
            set_dghot(d, n)
            run_ref(d)
            dump(d, fortio.read_case(d), name, n)


def storm_case():
    """NWS=8 Holland vortex over a 300 km x 200 km spherical basin with wetting and drying; the
    track mixes repeated BEST lines with an OFCL advisory (repeated increments included), so the
    run uses a hindcast-to-forecast interval with the nowcast's pressure and radii."""
    from dgswem_amd.holland import forecast_track
    case = mesh.storm_basin(3000, seed=3, nsteps=500, lx=3.0e5, ly=2.0e5, depth_range=(5.0, 200.0))
    case.params.dramp = 0.02
    case.storm.records = forecast_track(case.storm.records, issue=3)
    return case


def maxele_case(name: str, case, nsteps: int) -> None:
    """maxele.63 (ETAMAX, src/write_results.F:618-624) and the last fort.63 snapshot of a run whose
    global elevation output fires at its last step."""
    case.params.nt = nsteps
    with tempfile.TemporaryDirectory() as d:
        mesh.write_run_dir(d, case, dghot=0, dghotspool=nsteps, nspoolge=nsteps)
        run_ref(d)
        vals = np.loadtxt(os.path.join(d, "maxele.63"), skiprows=1)
        t, it, eta = fortio.read_fort63(os.path.join(d, "fort.63"))
        assert it[-1] == nsteps, it
        np.savez_compressed(os.path.join(HERE, f"{name}_maxele_step{nsteps}.npz"), etamax=vals,
                            eta=eta[-1], it=np.int64(nsteps))
        print(f"  {name}: maxele.63 after step {nsteps} written")


def delaunay4k():
    return mesh.delaunay_basin(n_elem=4000, seed=1, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 30.0),
                               cfl=0.3)


def main(which=None) -> None:
    if not os.path.exists(BIN):
        raise SystemExit(f"{BIN} missing: run oracle/build_ref.sh first")
    if which:   # a subset: python tests/golden/make_golden.py c1_8640 maxele
        if "c1_8640" in which:
            generated_case("c1", mesh.structured_basin(), [8640])
        if "maxele" in which:
            maxele_case("delaunay4k", delaunay4k(), 100)
        return
    generated_case("c1", mesh.structured_basin(), [1, 10, 1000, 8640])
    maxele_case("delaunay4k", delaunay4k(), 100)
    # linear friction (NOLIBF=0 needs a fort.15 TAU, so no Manning attribute), Coriolis, limiter and
    # a flux-specified side
    generated_case("c1var", mesh.structured_basin(nx=24, ny=12, slopeflag=5, cori=1.0e-4, nolibf=0,
                                                  manning=None, flow_side=True, flow_amp=2.0,
                                                  depth=8.0),
                   [200])
    generated_case("delaunay4k", delaunay4k(), [100, 1000])
    # wetting and drying without the slope limiter (WETDRY fused into the stage)
    generated_case("delaunay4k_nolim", mesh.delaunay_basin(n_elem=4000, seed=2, lx=6.0e4, ly=3.6e4,
                                                           depth_range=(2.0, 30.0), cfl=0.3,
                                                           slopeflag=0),
                   [300])
    generated_case("storm3k", storm_case(), [100, 500])
    reference_case("quarter_annular", [100, 2000])
    reference_case("wetdry", [50, 500])
    reference_case("mass_conservation", [1000])


if __name__ == "__main__":
    main(sys.argv[1:])
