"""The batched storm ensemble at the size the benchmark measures it: 64 spherical Delaunay basins of
20k-300k elements (log-uniform, 6.49M elements in all), each under its own Holland hurricane, with
the slope limiter and wetting/drying on, advanced together in one stage launch per RK stage with
graph replay, exactly as the benchmark builds them.

After 20 steps every member's state must match the oracle run of that member alone (relative L2
<= 1e-10 in elevation and velocity, identical wet/dry masks).
"""
import os

import numpy as np
import pytest

from dgswem_amd import gpu, mesh
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
STEPS = 20
THREADS = min(16, os.cpu_count() or 1)


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_full_size_storm_ensemble_matches_oracle():
    rng = np.random.default_rng(0)
    sizes = np.exp(rng.uniform(np.log(20_000), np.log(300_000), 64)).astype(int)
    cases = [mesh.storm_basin(int(n), seed=100 + i, nsteps=STEPS) for i, n in enumerate(sizes)]
    print(f"storm ensemble: {len(cases)} members, {sum(c.ne for c in cases)} elements, {STEPS} steps", flush=True)
    solvers = [gpu.GpuSolver(c) for c in cases]
    ens = gpu.Ensemble(solvers)
    for s in solvers:
        s.close()
    ens.use_graph(True)
    ens.step(STEPS)
    parts = ens.split(ens.state())
    ens.close()
    worst = 0.0
    for i, (c, a) in enumerate(zip(cases, parts)):
        o = Oracle(c, threads=THREADS)
        o.step(STEPS)
        b = o.state()
        np.testing.assert_array_equal(a[3], b[3])
        r = max(rel_l2(x, y) for x, y in zip(a[:3], b[:3]))
        worst = max(worst, r)
        assert np.isfinite(a[0]).all() and r <= TOL, (i, c.ne, r)
    print(f"worst member relative L2 {worst:.3e}", flush=True)
