"""The discrete adjoint at the size the benchmark measures it: the ~200k-element all-wet Delaunay
basin with Gaussian-random-field depth 10-3000 m (no limiter), 2000 steps with sqrt(N)
checkpointing, 50 gauges, misfit against a perturbed-Manning twin: the benchmark's adjoint
workload itself.  CF = 1e-4 keeps Manning's g n^2 / H^(1/3) above the friction floor at every
depth of the range (at 3000 m it is about 4e-4), so the loss depends on every node's n.

The gradient with respect to per-node Manning n and depth must match fourth-order central
differences of the same GPU forward model in 20 seeded directions to 1e-6 relative.
"""
import numpy as np
import pytest

from dgswem_amd import adjoint, mesh

pytestmark = pytest.mark.gpu

TOL = 1e-6
EPS = 1e-4
STEPS = 2000


def test_full_size_gradient_matches_central_differences():
    case = mesh.delaunay_basin(n_elem=200_000, seed=0, depth_range=(10.0, 3000.0), shelf=False,
                               wetdry=False, slopeflag=0, nsteps=STEPS, cf=1e-4)
    print(f"adjoint at full size: {case.ne} elements, {case.np_} nodes, {STEPS} steps", flush=True)
    prob = adjoint.twin_problem(case, nsteps=STEPS, ng=50, seed=0)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    assert J > 0.0 and np.all(np.isfinite(gn)) and np.all(np.isfinite(gdp))
    rng = np.random.default_rng(2024)
    for i in range(20):
        vn = rng.standard_normal(case.np_) * case.manning
        vd = rng.standard_normal(case.np_) * case.dp
        loss = lambda t: prob.loss(case.manning + t * vn, case.dp + t * vd)
        fd = (8.0 * (loss(EPS) - loss(-EPS)) - (loss(2 * EPS) - loss(-2 * EPS))) / (12.0 * EPS)
        ad = float(gn @ vn + gdp @ vd)
        rel = abs(fd - ad) / abs(ad)
        print(f"direction {i}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= TOL
    prob.solver.close()
