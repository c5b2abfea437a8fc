"""Batched discrete adjoint: an ensemble of p=2 storm basins, each under its own Holland vortex
(NWS=8) with its own gauges and perturbed-Manning twin data, slope limiter and wetting/drying on
with their active sets frozen from the forward run, differentiated in one batched forward and one
batched reverse sweep (the 32 x 100k p=2 batched-adjoint workload in miniature).

Checks: each member's slice of the batched gradient equals that member's own single-run gradient
bit for bit (the members are independent and every element's arithmetic is the same); the batched
gradient matches fourth-order central differences of the batched GPU forward model (relative
1e-6 in Manning-only directions, 1e-5 in directions that move the bathymetry, whose difference
step must stay below the limiter's branch switches); member 0's gradient matches central
differences of the CPU oracle in FP64 (the CPU reference gradient, relative 1e-6), and also
in a direction that moves the bathymetry when the oracle replays the forward run's limiter
branches (the frozen-branch map the adjoint differentiates, relative 1e-6).
"""
import copy

import numpy as np
import pytest

from dgswem_amd import adjoint, mesh
from oracle.oracle import Oracle, basis_at

pytestmark = pytest.mark.gpu

TOL = 1e-6
TOL_LIMITER = 1e-5
EPS = 1e-4
NSTEPS = 40


def fd4(loss, eps=EPS):
    return (8.0 * (loss(eps) - loss(-eps)) - (loss(2 * eps) - loss(-2 * eps))) / (12.0 * eps)


def members():
    """p=2 storm basins with the slope limiter and wetting/drying on (all wet: WETDRY's active set
    is the identity), each under its own Holland vortex."""
    return [mesh.storm_basin(ne, seed=seed, nsteps=NSTEPS, p=2, slopeflag=5, wetdry=True, shelf=False,
                             cfl=0.15, lx=3.0e5, ly=2.0e5, depth_range=(3.0, 20.0), cf=1e-4)
            for ne, seed in ((1500, 61), (2500, 62), (1200, 63))]


@pytest.fixture(scope="module")
def batch():
    cases = members()
    eprob, probs = adjoint.ensemble_problem(cases, nsteps=NSTEPS, ng=10, seed=70, state0="hump", perturb=0.5)
    mn = np.concatenate([c.manning for c in cases])
    dp = np.concatenate([c.dp for c in cases])
    J, gn, gdp = eprob.gradient(mn, dp)
    return cases, eprob, probs, mn, dp, J, gn, gdp


def test_members_equal_single_runs_bitwise(batch):
    cases, eprob, probs, mn, dp, J, gn, gdp = batch
    noff = eprob.solver.noff
    Js = 0.0
    for i, (c, p) in enumerate(zip(cases, probs)):
        Ji, gni, gdpi = p.gradient(c.manning, c.dp)
        Js += Ji
        print(f"member {i}: J {Ji:.6e}, |g_n| {np.abs(gni).max():.3e}", flush=True)
        assert np.abs(gni).max() > 0.0
        np.testing.assert_array_equal(gn[noff[i]:noff[i + 1]], gni)
        np.testing.assert_array_equal(gdp[noff[i]:noff[i + 1]], gdpi)
    assert J == pytest.approx(Js, rel=1e-12, abs=0.0)


def test_batched_gradient_matches_central_differences(batch):
    cases, eprob, probs, mn, dp, J, gn, gdp = batch
    rng = np.random.default_rng(9)
    for d in range(3):
        vn = rng.standard_normal(mn.size) * mn
        vd = rng.standard_normal(dp.size) * dp if d < 2 else np.zeros(dp.size)
        ad = float(gn @ vn + gdp @ vd)
        # depth steps of 1e-7 and more can reach a limiter branch switch (see test_gpu_adjoint.py);
        # Manning-only directions move the state little and take the larger, less noisy step
        fd = fd4(lambda t: eprob.loss(mn + t * vn, dp + t * vd), 1e-8 if d < 2 else EPS)
        print(f"direction {d}: adjoint {ad:.12e} fd {fd:.12e}", flush=True)
        assert abs(fd - ad) <= (TOL_LIMITER if d < 2 else TOL) * abs(ad)


def test_member_gradient_matches_cpu_oracle_differences(batch):
    """FP64 CPU reference: fourth-order central differences of the oracle's forward model of
    member 0 (its Holland forcing included) in the Manning direction."""
    cases, eprob, probs, mn, dp, J, gn, gdp = batch
    c0, p0 = cases[0], probs[0]
    elem, r, s = adjoint.seeded_gauges(c0, 10, 70)
    phi = np.stack([basis_at(c0.dg.pl, r[g], s[g]) for g in range(len(elem))])

    def oracle_loss(m):
        c = copy.copy(c0)
        c.manning = m
        o = Oracle(c, threads=8)
        o.set_state(*p0.state0)
        Jo = 0.0
        for k in range(NSTEPS):
            o.step(1)
            eta = np.sum(o.state()[0][elem - 1] * phi, axis=1)
            Jo += np.sum((eta - p0.data[k]) ** 2)
        return p0.weight * Jo

    j0 = oracle_loss(c0.manning)
    assert j0 == pytest.approx(p0.loss(c0.manning, c0.dp), rel=1e-8, abs=0.0)
    rng = np.random.default_rng(10)
    vn = rng.standard_normal(c0.np_) * c0.manning
    ad = float(gn[:c0.np_] @ vn)
    fd = fd4(lambda t: oracle_loss(c0.manning + t * vn))
    print(f"oracle fd {fd:.12e} adjoint {ad:.12e}", flush=True)
    assert abs(fd - ad) <= TOL * abs(ad)


def test_member_gradient_matches_frozen_branch_oracle(batch):
    """Member 0 against the map the adjoint differentiates, on the CPU: the oracle's forward model
    (Holland forcing included) with the slope limiter's branches replayed from the unperturbed
    run, fourth-order central differences at step 1e-5 in a direction that moves the bathymetry
    as well as Manning's n: relative 1e-6."""
    cases, eprob, probs, mn, dp, J, gn, gdp = batch
    c0, p0 = cases[0], probs[0]
    elem, r, s = adjoint.seeded_gauges(c0, 10, 70)
    phi = np.stack([basis_at(c0.dg.pl, r[g], s[g]) for g in range(len(elem))])
    o = Oracle(c0, threads=8)
    o.set_state(*p0.state0)
    o.limiter_record(NSTEPS)
    o.step(NSTEPS)
    records = o.limiter_records()
    assert np.mean((records[1] & 0x3F) != 0) > 0.001

    def frozen_loss(m, d):
        c = copy.copy(c0)
        c.manning = m
        c.dp = d
        oc = Oracle(c, threads=8)
        oc.set_state(*p0.state0)
        oc.limiter_replay(records)
        Jo = 0.0
        for k in range(NSTEPS):
            oc.step(1)
            eta = np.sum(oc.state()[0][elem - 1] * phi, axis=1)
            Jo += np.sum((eta - p0.data[k]) ** 2)
        return p0.weight * Jo

    rng = np.random.default_rng(12)
    n0 = c0.np_
    vn = rng.standard_normal(n0) * c0.manning
    vd = rng.standard_normal(n0) * c0.dp
    ad = float(gn[:n0] @ vn + gdp[:n0] @ vd)
    fd = fd4(lambda t: frozen_loss(c0.manning + t * vn, c0.dp + t * vd), 1e-5)
    rel = abs(fd - ad) / abs(ad)
    print(f"frozen-branch oracle fd {fd:.12e} adjoint {ad:.12e} rel {rel:.2e}", flush=True)
    assert rel <= TOL


def test_members_with_a_dry_coast_equal_single_runs_bitwise():
    """Members whose coast is dry (a shelf rising 2 m above the datum; from the cold start, so
    WETDRY acts every stage and each sweep runs the wet/dry reverse): each member's slice of the
    batched gradient equals the member's single-run gradient bit for bit."""
    cases = [mesh.storm_basin(ne, seed=seed, nsteps=NSTEPS, p=2, slopeflag=5, wetdry=True, shelf=True,
                              cfl=0.15, lx=3.0e5, ly=2.0e5, depth_range=(3.0, 20.0), cf=1e-4)
             for ne, seed in ((1500, 64), (2500, 65))]
    eprob, probs = adjoint.ensemble_problem(cases, nsteps=NSTEPS, ng=10, seed=80, perturb=0.5)
    for p in probs:
        assert (p.state0[3] == 0).any()            # the cold start dried part of the coast
    mn = np.concatenate([c.manning for c in cases])
    dp = np.concatenate([c.dp for c in cases])
    J, gn, gdp = eprob.gradient(mn, dp)
    noff = eprob.solver.noff
    for i, (c, p) in enumerate(zip(cases, probs)):
        Ji, gni, gdpi = p.gradient(c.manning, c.dp)
        print(f"dry-coast member {i}: J {Ji:.6e}, |g_dp| {np.abs(gdpi).max():.3e}", flush=True)
        assert np.abs(gdpi).max() > 0.0
        np.testing.assert_array_equal(gn[noff[i]:noff[i + 1]], gni)
        np.testing.assert_array_equal(gdp[noff[i]:noff[i + 1]], gdpi)
        p.solver.close()
    eprob.solver.close()
