"""Long horizons at the benchmark sizes, against the oracle (OpenMP build), with the relative L2
difference printed as it develops.  SLOPELIMITER5 answers last-bit differences discontinuously, so
these are the runs where a difference in rounding would grow; the GPU evaluates the step's
transcendental terms correctly rounded and every other operation in the reference's order, and
the states stay within round-off (about 3e-14 after 1000 steps at p=1).

* the 997,306-element tidal basin (limiter 5, wetting/drying, five constituents): 1000 steps at
  p=1, 200 steps at p=2;
* eight members of the 64-member storm ensemble (the benchmark's first eight sizes), batched with
  graph replay: 200 steps, or all 1000 of the ensemble configuration with DGSWEM_FULL_CHECKS=1
  (about two minutes of oracle time more).
Bar: relative L2 <= 1e-10 in elevation and velocity and identical wet/dry masks at every check.
"""
import os

import numpy as np
import pytest

from dgswem_amd import gpu, mesh
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
THREADS = min(16, os.cpu_count() or 1)


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("p,steps,chunk", [(1, 1000, 250), (2, 200, 50)])
def test_full_size_long_horizon(p, steps, chunk):
    case = mesh.delaunay_basin(n_elem=1_000_000, seed=0, p=p, cfl=0.3 if p == 1 else 0.15,
                               nsteps=steps)
    print(f"full-size basin p={p}: {case.ne} elements, {steps} steps", flush=True)
    g = gpu.GpuSolver(case)
    o = Oracle(case, threads=THREADS)
    worst = 0.0
    for k in range(chunk, steps + 1, chunk):
        g.step(chunk)
        o.step(chunk)
        a, b = g.state(), o.state()
        r = max(rel_l2(x, y) for x, y in zip(a[:3], b[:3]))
        worst = max(worst, r)
        print(f"  step {k}: rel L2 {r:.3e}, masks equal {bool(np.array_equal(a[3], b[3]))}", flush=True)
        np.testing.assert_array_equal(a[3], b[3])
        assert np.isfinite(a[0]).all() and r <= TOL
    g.close()


def test_storm_members_long_horizon():
    rng = np.random.default_rng(0)
    sizes = np.exp(rng.uniform(np.log(20_000), np.log(300_000), 64)).astype(int)[:8]
    steps, chunk = (1000, 250) if os.environ.get("DGSWEM_FULL_CHECKS") == "1" else (200, 50)
    cases = [mesh.storm_basin(int(n), seed=100 + i, nsteps=steps) for i, n in enumerate(sizes)]
    print(f"storm members: {len(cases)}, {sum(c.ne for c in cases)} elements, {steps} steps", flush=True)
    solvers = [gpu.GpuSolver(c) for c in cases]
    ens = gpu.Ensemble(solvers)
    for s in solvers:
        s.close()
    ens.use_graph(True)
    oracles = [Oracle(c, threads=THREADS) for c in cases]
    for k in range(chunk, steps + 1, chunk):
        ens.step(chunk)
        parts = ens.split(ens.state())
        worst = 0.0
        for i, (o, a) in enumerate(zip(oracles, parts)):
            o.step(chunk)
            b = o.state()
            r = max(rel_l2(x, y) for x, y in zip(a[:3], b[:3]))
            worst = max(worst, r)
            np.testing.assert_array_equal(a[3], b[3])
            assert np.isfinite(a[0]).all() and r <= TOL, (k, i, r)
        print(f"  step {k}: worst member rel L2 {worst:.3e}", flush=True)
    ens.close()
