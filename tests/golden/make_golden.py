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
            if name.startswith(("delaunay", "storm", "tipot", "c1eseries", "c1fseries")) and n == steps[0]:
                keep = os.path.join(HERE, name)
                os.makedirs(keep, exist_ok=True)
                for f in ("fort.14", "fort.15", "fort.13", "fort.dg", "fort.22", "fort.24", "fort.19",
                          "fort.20"):
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


# Tidal potential constituents (TPK, AMIGT rad/s, ETRF): the usual ADCIRC fort.15 values
TIDAL_POTENTIAL = [("M2", 0.242334, 1.405189e-4, 0.693), ("S2", 0.112841, 1.454441e-4, 0.693),
                   ("N2", 0.046398, 1.378797e-4, 0.693), ("K1", 0.141565, 7.292117e-5, 0.736),
                   ("O1", 0.100514, 6.759774e-5, 0.695)]


def tidal_potential_case(ntip: int, nsteps: int):
    """A 2000-element spherical basin (1000 km x 600 km around 28 N, limiter and wetting/drying on)
    under the five tidal potential constituents (two diurnal, three semidiurnal) besides its open
    boundary tides; with ntip=2 also seeded self-attraction/load amplitudes and phases."""
    case = mesh.delaunay_basin(n_elem=2000, seed=11, ics=2, nsteps=nsteps, depth_range=(5.0, 3000.0))
    case.params.dramp = 0.05
    p = case.params
    rng = np.random.default_rng(12)
    p.ntip = ntip
    p.ntif = len(TIDAL_POTENTIAL)
    p.tipotag = [t[0] for t in TIDAL_POTENTIAL]
    p.tpk = np.array([t[1] for t in TIDAL_POTENTIAL])
    p.amigt = np.array([t[2] for t in TIDAL_POTENTIAL])
    p.etrf = np.array([t[3] for t in TIDAL_POTENTIAL])
    p.fft = rng.uniform(0.9, 1.1, p.ntif)
    p.facet = rng.uniform(0.0, 360.0, p.ntif) * fortio.DEG2RAD
    p.pert = np.array([2.0 * fortio.PI / a for a in p.amigt])
    if ntip == 2:
        lat = case.sfea / fortio.DEG2RAD
        p.saltamp = np.stack([0.01 + 0.02 * np.abs(np.sin(lat * (k + 1))) for k in range(p.ntif)])
        p.saltpha = np.stack([(30.0 * k + 10.0 * np.cos(case.slam * 50.0)) for k in range(p.ntif)]) * fortio.DEG2RAD
    return case


def elev_series_case(nsteps: int):
    """The config-1 basin with its open boundary driven by a fort.19 time series (NBFR = 0):
    hourly records of a semidiurnal-like signal that varies along the boundary."""
    case = mesh.structured_basin(nsteps=nsteps)
    p = case.params
    p.nbfr = 0
    p.amig = p.ff = p.face = p.per = np.zeros(0)
    p.emo = p.efa = np.zeros((0, 0))
    neta = len(case.nbd())
    t = np.arange(40) * 3600.0
    along = np.linspace(0.0, 1.0, neta)
    vals = 0.4 * np.sin(2 * np.pi * t[:, None] / 44712.0 + 0.3 * along[None, :]) + 0.05 * along[None, :]
    case.elev_series = (3600.0, vals)
    return case


def flow_series_case(nsteps: int):
    """The c1var basin (limiter, Coriolis, linear friction) with its flux side driven by a fort.20
    time series (NFFR = 0): half-hourly records of an inflow that varies along the side."""
    case = mesh.structured_basin(nx=24, ny=12, slopeflag=5, cori=1.0e-4, nolibf=0, manning=None,
                                 flow_side=True, flow_amp=2.0, depth=8.0, nsteps=nsteps)
    p = case.params
    p.nffr = 0
    p.famig = p.fff = p.fface = p.fper = np.zeros(0)
    p.qnam = p.qnph = np.zeros((1, 0))
    nflux = sum(len(n) for tp, n in case.land_segs if tp in fortio.FLOW_TYPES)
    t = np.arange(30) * 1800.0
    along = np.linspace(0.0, 1.0, nflux)
    vals = -1.5 * np.sin(2 * np.pi * t[:, None] / 30000.0) * (1.0 + 0.2 * along[None, :])
    case.flow_series = (1800.0, vals)
    return case


def delaunay4k():
    return mesh.delaunay_basin(n_elem=4000, seed=1, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 30.0),
                               cfl=0.3)


def forcing_cases() -> None:
    """Tidal potential (NTIP = 1 and 2) and the non-periodic boundaries (fort.19, fort.20)."""
    generated_case("tipot2k", tidal_potential_case(1, 300), [300])
    generated_case("tipotsal2k", tidal_potential_case(2, 200), [200])
    generated_case("c1eseries", elev_series_case(1000), [1000])
    generated_case("c1fseries", flow_series_case(400), [400])


def main(which=None) -> None:
    if not os.path.exists(BIN):
        raise SystemExit(f"{BIN} missing: run oracle/build_ref.sh first")
    if which:   # a subset: python tests/golden/make_golden.py c1_8640 maxele
        if "c1_8640" in which:
            generated_case("c1", mesh.structured_basin(), [8640])
        if "maxele" in which:
            maxele_case("delaunay4k", delaunay4k(), 100)
        if "forcing" in which:
            forcing_cases()
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
    forcing_cases()
    reference_case("quarter_annular", [100, 2000])
    reference_case("wetdry", [50, 500])
    reference_case("mass_conservation", [1000])


if __name__ == "__main__":
    main(sys.argv[1:])
