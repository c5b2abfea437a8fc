"""The batched adjoint at the size the benchmark measures it: 32 spherical storm basins of 100k
elements at p=2 (3.17M elements), slope limiter 5 and wetting/drying on (all wet) with their
active sets frozen from the forward run, each member under its own Holland vortex with 20 gauges
and a perturbed-Manning twin, 1000 steps with checkpoint spacing 32: the benchmark's
batched-adjoint workload itself.

* The batched gradient against fourth-order central differences of the batched GPU forward model
  in three seeded Manning directions, relative 1e-5 at difference step 1e-5.  The forward model
  differenced here is the real map, whose slope-limiter branches (and the LLF wave-speed maximum)
  switch under the perturbation while the adjoint keeps the forward run's branches: on this
  workload the two agree to 1.4e-6 .. 5.1e-6 at step 1e-5; a step of 1e-4 crosses more switches
  (3e-6 .. 1e-2 with a moving hump), 1e-6 is dominated by rounding (2e-5).
* The map the adjoint differentiates, at this workload's mesh size and order: one member's
  single-run gradient over 100 steps from a Gaussian hump, against the oracle (FP64, CPU) with the
  slope limiter's branches replayed from the unperturbed run, fourth-order differences at step
  1e-5 in a direction that moves Manning's n and the bathymetry: relative 1e-6.
"""
import copy
import os

import numpy as np
import pytest

from dgswem_amd import adjoint, mesh
from oracle.oracle import Oracle, basis_at

pytestmark = pytest.mark.gpu

STEPS = 1000
MEMBERS = 32
TOL_REAL_MAP = 1e-5
TOL_FROZEN = 1e-6
EPS = 1e-5
THREADS = min(16, os.cpu_count() or 1)


def fd4(loss, eps=EPS):
    return (8.0 * (loss(eps) - loss(-eps)) - (loss(2 * eps) - loss(-2 * eps))) / (12.0 * eps)


def member_case(i, nsteps=STEPS):
    return mesh.storm_basin(100_000, seed=300 + i, nsteps=nsteps, p=2, slopeflag=5, wetdry=True,
                            shelf=False, cfl=0.15, depth_range=(10.0, 300.0), cf=1e-4)


@pytest.fixture(scope="module")
def batch():
    cases = [member_case(i) for i in range(MEMBERS)]
    print(f"batched adjoint at full size: {MEMBERS} members, {sum(c.ne for c in cases)} elements, p=2, "
          f"{STEPS} steps", flush=True)
    eprob, probs = adjoint.ensemble_problem(cases, nsteps=STEPS, ng=20, seed=400)
    for p in probs:
        p.solver.close()
    mn = np.concatenate([c.manning for c in cases])
    dp = np.concatenate([c.dp for c in cases])
    eprob.interval = 32
    J, gn, gdp = eprob.gradient(mn, dp)
    print(f"J {J:.6e}", flush=True)
    yield eprob, mn, dp, J, gn, gdp
    eprob.solver.close()


def test_full_size_gradient_is_finite(batch):
    eprob, mn, dp, J, gn, gdp = batch
    assert J > 0.0 and np.isfinite(gn).all() and np.isfinite(gdp).all()
    assert np.abs(gn).max() > 0.0


@pytest.mark.parametrize("d", [0, 1, 2])
def test_full_size_gradient_matches_central_differences(batch, d):
    eprob, mn, dp, J, gn, gdp = batch
    vn = np.random.default_rng(2025 + d).standard_normal(mn.size) * mn
    fd = fd4(lambda t: eprob.loss(mn + t * vn, dp))
    ad = float(gn @ vn)
    rel = abs(fd - ad) / abs(ad)
    print(f"direction {d}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
    assert rel <= TOL_REAL_MAP


def test_member_gradient_matches_frozen_branch_oracle():
    nsteps = 100
    c0 = member_case(0, nsteps)
    st0 = adjoint.hump_state(c0)
    p0 = adjoint.twin_problem(c0, nsteps=nsteps, ng=20, seed=400, perturb=0.5, state0=st0)
    J0, gn, gdp = p0.gradient(c0.manning, c0.dp)
    p0.solver.close()
    elem, r, s = adjoint.seeded_gauges(c0, 20, 400)
    phi = np.stack([basis_at(c0.dg.pl, r[g], s[g]) for g in range(len(elem))])
    o = Oracle(c0, threads=THREADS)
    o.set_state(*st0)
    o.limiter_record(nsteps)
    o.step(nsteps)
    records = o.limiter_records()
    assert np.mean((records[1] & 0x3F) != 0) > 0.001

    def frozen_loss(m, dpt):
        c = copy.copy(c0)
        c.manning = m
        c.dp = dpt
        oc = Oracle(c, threads=THREADS)
        oc.set_state(*st0)
        oc.limiter_replay(records)
        Jo = 0.0
        for k in range(nsteps):
            oc.step(1)
            eta = np.sum(oc.state()[0][elem - 1] * phi, axis=1)
            Jo += np.sum((eta - p0.data[k]) ** 2)
        return p0.weight * Jo

    assert frozen_loss(c0.manning, c0.dp) == pytest.approx(J0, rel=1e-10)
    rng = np.random.default_rng(12)
    vn = rng.standard_normal(c0.np_) * c0.manning
    vd = rng.standard_normal(c0.np_) * c0.dp
    ad = float(gn @ vn + gdp @ vd)
    fd = fd4(lambda t: frozen_loss(c0.manning + t * vn, c0.dp + t * vd))
    rel = abs(fd - ad) / abs(ad)
    print(f"frozen-branch oracle, 100k p=2: fd {fd:.12e} adjoint {ad:.12e} rel {rel:.2e}", flush=True)
    assert rel <= TOL_FROZEN
