"""Discrete adjoint of the GPU step (through the C ABI and the torch front end).

Gradients of the gauge-misfit loss with respect to per-node Manning n and depth DP are checked
against fourth-order central finite differences of the same GPU forward model in 20 seeded
directions (relative agreement 1e-6), against central differences of the CPU oracle (the FP64 CPU reference
gradient), and for independence from the checkpoint spacing (bit for bit).
"""
import numpy as np
import pytest

from dgswem_amd import adjoint, gpu, mesh

pytestmark = pytest.mark.gpu

TOL = 1e-6
TOL_LIMITER = 1e-5      # slope limiter on: rounding-limited difference steps (see the limiter test)


def small_case(p=1, n_elem=2000, seed=31, nsteps=60):
    """All-wet seeded Delaunay basin, no limiter: the differentiated configuration.  Shallow (2-10 m)
    with CF=1e-4, so that Manning friction is above its floor at every node."""
    return mesh.delaunay_basin(n_elem=n_elem, seed=seed, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 10.0),
                               shelf=False, wetdry=False, slopeflag=0, p=p, nsteps=nsteps, cf=1e-4)


def directions(case, k, seed):
    rng = np.random.default_rng(seed)
    for _ in range(k):
        vn = rng.standard_normal(case.np_) * case.manning
        vd = rng.standard_normal(case.np_) * case.dp
        yield vn, vd


EPS = 1e-4


def fd4(loss, eps=EPS):
    """Fourth-order central difference of t -> loss(t) at 0: truncation O(eps^4), so eps can be
    large enough that the forward model's rounding stays far below the tolerance."""
    return (8.0 * (loss(eps) - loss(-eps)) - (loss(2 * eps) - loss(-2 * eps))) / (12.0 * eps)


def central_fd(problem, case, vn, vd, eps=EPS):
    return fd4(lambda t: problem.loss(case.manning + t * vn, case.dp + t * vd), eps)


@pytest.fixture(scope="module")
def p1():
    case = small_case()
    prob = adjoint.twin_problem(case, nsteps=60, ng=20, seed=5)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    return case, prob, J, gn, gdp


def test_loss_matches_forward(p1):
    case, prob, J, gn, gdp = p1
    print(f"J = {J:.6e}", flush=True)
    assert J > 0.0
    assert J == pytest.approx(prob.loss(case.manning, case.dp), rel=1e-13, abs=0.0)
    assert np.all(np.isfinite(gn)) and np.all(np.isfinite(gdp))
    assert np.abs(gn).max() > 0.0 and np.abs(gdp).max() > 0.0


def test_gradient_matches_central_differences_20_directions(p1):
    case, prob, J, gn, gdp = p1
    worst = 0.0
    for i, (vn, vd) in enumerate(directions(case, 20, seed=77)):
        ad = float(gn @ vn + gdp @ vd)
        fd = central_fd(prob, case, vn, vd)
        rel = abs(fd - ad) / max(abs(ad), 1e-300)
        worst = max(worst, rel)
        print(f"direction {i}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= TOL
    print(f"worst relative difference {worst:.2e}", flush=True)


def test_manning_and_depth_parts_separately(p1):
    case, prob, J, gn, gdp = p1
    rng = np.random.default_rng(3)
    vn = rng.standard_normal(case.np_) * case.manning
    vd = rng.standard_normal(case.np_) * case.dp
    z = np.zeros(case.np_)
    for a, b, g in ((vn, z, gn @ vn), (z, vd, gdp @ vd)):
        fd = central_fd(prob, case, a, b)
        assert abs(fd - g) <= TOL * abs(g)


def test_checkpoint_spacing_does_not_change_gradient(p1):
    case, prob, J, gn, gdp = p1
    for interval in (1, 7, 60):
        prob.interval = interval
        J2, gn2, gdp2 = prob.gradient(case.manning, case.dp)
        assert J2 == J
        np.testing.assert_array_equal(gn2, gn)
        np.testing.assert_array_equal(gdp2, gdp)
    prob.interval = 0


def test_gradient_leaves_state_unchanged(p1):
    case, prob, J, gn, gdp = p1
    prob.reset()
    s0 = prob.solver.state()
    prob.solver.gradient(prob.data, prob.weight)
    for a, b in zip(prob.solver.state(), s0):
        np.testing.assert_array_equal(a, b)


def test_torch_autograd_front_end(p1):
    torch = pytest.importorskip("torch")
    case, prob, J, gn, gdp = p1
    m = torch.tensor(case.manning, dtype=torch.float64, requires_grad=True)
    d = torch.tensor(case.dp, dtype=torch.float64, requires_grad=True)
    # one forward-and-adjoint call serves both forward and backward (no second forward run)
    calls = []
    orig_g, orig_l = prob.gradient, prob.loss
    prob.gradient = lambda *a, **k: (calls.append("gradient"), orig_g(*a, **k))[1]
    prob.loss = lambda *a, **k: (calls.append("loss"), orig_l(*a, **k))[1]
    try:
        loss = adjoint.gauge_misfit(m, d, prob)
        assert float(loss.detach()) == pytest.approx(J, rel=1e-13, abs=0.0)
        (2.0 * loss).backward()
        assert calls == ["gradient"]
        with torch.no_grad():
            plain = adjoint.gauge_misfit(m, d, prob)
        assert calls == ["gradient", "loss"]
    finally:
        prob.gradient, prob.loss = orig_g, orig_l
    assert float(plain) == pytest.approx(J, rel=1e-13, abs=0.0)
    np.testing.assert_array_equal(m.grad.numpy(), 2.0 * gn)
    np.testing.assert_array_equal(d.grad.numpy(), 2.0 * gdp)


def test_gradient_matches_cpu_oracle_differences(p1):
    """FP64 CPU reference: central differences of the oracle's forward model (the restatement of
    dgswem_serial's step) in three directions."""
    from oracle.oracle import Oracle
    case, prob, J, gn, gdp = p1
    elem, r, s = adjoint.seeded_gauges(case, 20, 5)
    phi = np.stack([np.ones_like(r), (1.0 + 3.0 * s) / 2.0, (1.0 + 2.0 * r + s) / 2.0], axis=1)

    ze0, qx0, qy0, wd0 = prob.state0

    def oracle_loss(mn, dp):
        import copy
        c = copy.copy(case)
        c.manning = mn; c.dp = dp
        o = Oracle(c, threads=8)
        o.set_state(ze0, qx0, qy0, wd0)
        J = 0.0
        for k in range(prob.nsteps):
            o.step(1)
            ze = o.state()[0]
            eta = np.sum(ze[elem - 1] * phi, axis=1)
            J += np.sum((eta - prob.data[k]) ** 2)
        return prob.weight * J

    j0 = oracle_loss(case.manning, case.dp)
    assert j0 == pytest.approx(J, rel=1e-9, abs=0.0)
    for vn, vd in directions(case, 3, seed=11):
        ad = float(gn @ vn + gdp @ vd)
        fd = fd4(lambda t: oracle_loss(case.manning + t * vn, case.dp + t * vd))
        print(f"oracle fd {fd:.12e} adjoint {ad:.12e}", flush=True)
        assert abs(fd - ad) <= TOL * abs(ad)


def test_p2_gradient_matches_central_differences():
    case = small_case(p=2, n_elem=1200, seed=32, nsteps=30)
    case.params.dtdp = case.params.dtdp * 0.5     # p=2 needs a smaller CFL number
    prob = adjoint.twin_problem(case, nsteps=30, ng=10, seed=6)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    for vn, vd in directions(case, 4, seed=12):
        ad = float(gn @ vn + gdp @ vd)
        fd = central_fd(prob, case, vn, vd)
        print(f"p2 adjoint {ad:.12e} fd {fd:.12e}", flush=True)
        assert abs(fd - ad) <= TOL * abs(ad)


def test_adjoint_rejects_runs_whose_wetting_and_drying_acts():
    """Without the slope limiter WETDRY runs inside the stage kernel and is differentiated only as
    the identity; a run in which it dries, rewets or redistributes elements is refused after the
    forward sweep, with the handle's state restored."""
    case = mesh.delaunay_basin(n_elem=1500, seed=33, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 30.0),
                               nsteps=10, slopeflag=0)   # wetting/drying on, dry land, no limiter
    s = gpu.GpuSolver(case)
    s.set_gauges(np.array([1, 2], np.int32), np.array([-0.5, -0.5]), np.array([-0.5, -0.5]))
    before = s.state()
    with pytest.raises(gpu.GpuError, match="wetting/drying changed"):
        s.gradient(np.zeros((5, 2)))
    for a, b in zip(s.state(), before):
        np.testing.assert_array_equal(a, b)
    s.close()


def test_adjoint_refusal_restores_the_wet_dry_flags():
    """A closed basin whose shoreline moves, without the slope limiter: WETDRY changes elements'
    flags during the 60-step window (the oracle run below shows it).  The refused gradient call
    must leave both the modal state and the wet/dry flags as they were, so the next step continues
    from there."""
    from oracle.oracle import Oracle
    from test_gpu_edge_cases import closed, wetdry_state
    case = closed(12, 6, 1, nolifa=2, slopeflag=0)
    case.dp = 2.0 - 0.0004 * case.x
    st = wetdry_state(case, 2.0)
    o = Oracle(case, threads=4)
    o.set_state(*st)
    o.step(60)
    assert (o.state()[3] != st[3]).any()       # the window does change a flag
    s = gpu.GpuSolver(case)
    s.set_state(*st)
    s.set_gauges(np.array([1, 40], np.int32), np.array([-0.5, -0.5]), np.array([-0.5, -0.5]))
    with pytest.raises(gpu.GpuError, match="wetting/drying changed"):
        s.gradient(np.zeros((60, 2)))
    for a, b in zip(s.state(), st):
        np.testing.assert_array_equal(a, b)
    # and stepping on from the restored state equals the oracle from the same state
    s.step(60)
    for a, b in zip(s.state(), o.state()):
        np.testing.assert_array_equal(a, b)
    s.close()


def limiter_case(p=1, wetdry=False, n_elem=2000, seed=34, nsteps=60):
    """All-wet shallow basin with the slope limiter on (and WETDRY on but inactive when wetdry)."""
    return mesh.delaunay_basin(n_elem=n_elem, seed=seed, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 10.0),
                               shelf=False, wetdry=wetdry, slopeflag=5, p=p, nsteps=nsteps, cf=1e-4)


@pytest.mark.parametrize("p,wetdry", [(1, False), (1, True), (2, True)])
def test_frozen_limiter_gradient_matches_central_differences(p, wetdry):
    """Slope limiter on: the reverse takes the forward's branches (clamps, redistribution tests,
    min(redfac, redmax)) as fixed; central differences of the true forward agree as long as the
    perturbation crosses no branch switch.  SLOPELIMITER5 is not even continuous (its
    redistribution count and sign tests jump), and steps of 1e-7..1e-5 (relative) in bathymetry
    reach a switch somewhere on these runs (a line scan of J shows the jump), so the step here is
    1e-8.  At that step FP64 rounding of the loss limits the agreement to about 1e-6, so the
    tolerance here is TOL_LIMITER; without the limiter the tests above hold 1e-6 at step 1e-4."""
    nst = 40 if p == 1 else 60
    case = limiter_case(p=p, wetdry=wetdry, nsteps=nst)
    if p == 2:
        case.params.dtdp = case.params.dtdp * 0.5
    st = adjoint.hump_state(case, amplitude=0.8, width=0.15)
    # the limiter does act on this run
    free = mesh.delaunay_basin(n_elem=2000, seed=34, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 10.0),
                               shelf=False, wetdry=wetdry, slopeflag=0, p=p, nsteps=nst, cf=1e-4)
    free.params.dtdp = case.params.dtdp
    a_, b_ = gpu.GpuSolver(case), gpu.GpuSolver(free)
    for g in (a_, b_):
        g.set_state(*st)
        g.step(10)
    assert not np.array_equal(a_.state()[1], b_.state()[1])
    a_.close(); b_.close()
    prob = adjoint.twin_problem(case, nsteps=nst, ng=40, seed=8, state0=st, perturb=0.5)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    assert np.abs(gn).max() > 0.0
    for vn, vd in directions(case, 3, seed=13):
        ad = float(gn @ vn + gdp @ vd)
        fd = central_fd(prob, case, vn, vd, eps=1e-8)
        print(f"p={p} wetdry={wetdry}: adjoint {ad:.12e} fd {fd:.12e} rel {abs(fd - ad) / abs(ad):.2e}", flush=True)
        assert abs(fd - ad) <= TOL_LIMITER * abs(ad)


def frozen_oracle_loss(case, prob, ng, seed, records):
    """Loss of the oracle's forward model with the slope limiter's branches replayed from `records`
    (the frozen-branch map the adjoint differentiates): a function of Manning n and depth."""
    import copy
    from oracle.oracle import Oracle, basis_at
    elem, r, s = adjoint.seeded_gauges(case, ng, seed)
    phi = np.stack([basis_at(case.dg.pl, r[g], s[g]) for g in range(len(elem))])

    def loss(mn, dp):
        c = copy.copy(case)
        c.manning = mn
        c.dp = dp
        o = Oracle(c, threads=8)
        o.set_state(*prob.state0)
        o.limiter_replay(records)
        J = 0.0
        for k in range(prob.nsteps):
            o.step(1)
            eta = np.sum(o.state()[0][elem - 1] * phi, axis=1)
            J += np.sum((eta - prob.data[k]) ** 2)
        return prob.weight * J
    return loss


@pytest.mark.parametrize("p,wetdry", [(1, False), (1, True), (2, True)])
def test_frozen_limiter_gradient_matches_frozen_branch_oracle(p, wetdry):
    """The map the adjoint differentiates, evaluated on the CPU: the oracle's forward model with
    every SLOPELIMITER5 branch (the clamped vertices and the bound-attaining elements, the signs,
    counts and min choices of the redistribution passes) replayed from the unperturbed run.  Its
    fourth-order central differences, in directions that move the bathymetry as well as Manning's
    n, match the GPU gradient to 1e-6.  The step is 1e-5: the map is smooth but for the LLF wave
    speed bound, a maximum that is continuous, whose kinks give the differences an error that
    shrinks in proportion to the step (on a run of this basin the differences at steps 1e-3, 1e-4
    and 1e-5 moved by 2e-7 and 2e-8 relative)."""
    from oracle.oracle import Oracle
    nst = 40 if p == 1 else 60
    case = limiter_case(p=p, wetdry=wetdry, nsteps=nst)
    if p == 2:
        case.params.dtdp = case.params.dtdp * 0.5
    st = adjoint.hump_state(case, amplitude=0.8, width=0.15)
    prob = adjoint.twin_problem(case, nsteps=nst, ng=40, seed=8, state0=st, perturb=0.5)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    o = Oracle(case, threads=8)
    o.set_state(*prob.state0)
    o.limiter_record(nst)
    o.step(nst)
    records = o.limiter_records()
    assert np.mean((records[1] & 0x3F) != 0) > 0.01        # the limiter clamps on this run
    loss = frozen_oracle_loss(case, prob, 40, 8, records)
    assert loss(case.manning, case.dp) == pytest.approx(J, rel=1e-9, abs=0.0)
    for vn, vd in directions(case, 3, seed=13):
        ad = float(gn @ vn + gdp @ vd)
        fd = fd4(lambda t: loss(case.manning + t * vn, case.dp + t * vd), 1e-5)
        rel = abs(fd - ad) / abs(ad)
        print(f"p={p} wetdry={wetdry}: adjoint {ad:.12e} frozen-branch oracle fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= TOL
    prob.solver.close()


@pytest.mark.parametrize("p", [1, 2])
def test_active_wetting_and_drying_gradient_matches_frozen_branch_oracle(p):
    """Wetting and drying that acts, with the slope limiter on: a closed basin whose beach rises
    from 2 m deep to 2.8 m above the datum under a 2 m hump, so that every stage WETDRY keeps dry
    elements flat, redistributes the water of partly wet ones and dries or floods some, and the
    edges between wet and dry elements take the dry-side fluxes (reflection, zero gravity).  The
    reverse takes WETDRY's branches and the dry-side cases from the forward run; the oracle
    replays the recorded SLOPELIMITER5 and WETDRY branches (the frozen-branch map).  Its
    fourth-order differences at step 1e-6, in directions that move the bathymetry as well as
    Manning's n, match the GPU gradient to 1e-6.  The step is smaller than in the tests above
    because this map has more kinks (the LLF wave-speed maximum on every wet/dry edge), whose
    error shrinks in proportion to the step: at p=2, in the second direction (a derivative 100
    times smaller than the others), the differences at steps 1e-3, 1e-4, 1e-5, 1e-6 and 1e-7
    differed from the adjoint by 1.3e-3, 7.2e-5, 1.9e-5, 7.6e-7 and 1.2e-7 relative, while at p=1
    the steps 1e-5 and 1e-6 agree to 6e-8."""
    from oracle.oracle import Oracle
    from test_gpu_edge_cases import closed, wetdry_state
    nst = 60
    case = closed(12, 6, p, nolifa=2, slopeflag=5)
    case.dp = 2.0 - 0.0004 * case.x
    case.manning = 0.02 + 0.01 * np.random.default_rng(4).random(case.np_)
    st = wetdry_state(case, 2.0)
    prob = adjoint.twin_problem(case, nsteps=nst, ng=20, seed=8, state0=st, perturb=0.5)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    assert np.abs(gn).max() > 0.0 and np.abs(gdp).max() > 0.0
    # the gradient call leaves the handle where it was
    for a, b in zip(prob.solver.state(), st):
        np.testing.assert_array_equal(a, b)
    o = Oracle(case, threads=8)
    o.set_state(*st)
    o.limiter_record(nst)
    o.step(nst)
    records = o.limiter_records()
    kinds = np.bincount(records[2] & 3, minlength=4)
    print(f"p={p}: WETDRY branches identity/check/flat/redistribution {kinds.tolist()}", flush=True)
    assert kinds[2] > 0 and kinds[3] > 0
    loss = frozen_oracle_loss(case, prob, 20, 8, records)
    assert loss(case.manning, case.dp) == pytest.approx(J, rel=1e-9, abs=0.0)
    for vn, vd in directions(case, 3, seed=13):
        ad = float(gn @ vn + gdp @ vd)
        fd = fd4(lambda t: loss(case.manning + t * vn, case.dp + t * vd), 1e-6)
        rel = abs(fd - ad) / abs(ad)
        print(f"p={p}: adjoint {ad:.12e} frozen-branch oracle fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= TOL
    prob.solver.close()


def test_active_wetting_and_drying_gradient_independent_of_checkpoint_spacing():
    """The wet/dry flags the reverse sweep reads are kept per checkpoint and per stage of each
    recomputed segment; the gradient must not depend on where the segments start (spacings that
    divide the 60 steps and ones that do not), bit for bit."""
    from test_gpu_edge_cases import closed, wetdry_state
    case = closed(12, 6, 1, nolifa=2, slopeflag=5)
    case.dp = 2.0 - 0.0004 * case.x
    case.manning = 0.02 + 0.01 * np.random.default_rng(4).random(case.np_)
    st = wetdry_state(case, 2.0)
    prob = adjoint.twin_problem(case, nsteps=60, ng=20, seed=8, state0=st, perturb=0.5)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    for interval in (1, 7, 13, 60):
        prob.interval = interval
        J2, gn2, gdp2 = prob.gradient(case.manning, case.dp)
        assert J2 == J
        np.testing.assert_array_equal(gn2, gn)
        np.testing.assert_array_equal(gdp2, gdp)
    prob.solver.close()


@pytest.mark.parametrize("p", [1, 2])
def test_storm_with_a_dry_coast_gradient_matches_frozen_branch_oracle(p):
    """A spherical storm basin (NWS=8 Holland forcing) whose coast the cold start dries, with a
    hump added on the wet elements: the wet/dry variant of the spherical Holland adjoint kernel,
    against the frozen-branch oracle (SLOPELIMITER5, WETDRY and the dry-side edge cases replayed).
    The loss is the gauges' mean square elevation (data zero), so that last-bit differences of
    the Holland terms between the GPU and the oracle stay far below the misfit."""
    from oracle.oracle import Oracle
    nst = 60
    case = mesh.storm_basin(2000, seed=66, nsteps=nst, p=p, slopeflag=5, wetdry=True, shelf=True,
                            cfl=0.15 if p == 2 else 0.3, lx=3.0e5, ly=2.0e5, depth_range=(3.0, 20.0), cf=1e-4)
    g = gpu.GpuSolver(case)
    cold = g.state()
    g.close()
    assert (cold[3] == 0).any()
    hump = adjoint.hump_state(case, amplitude=0.5)
    ze = cold[0].copy()
    wet = cold[3] == 1
    ze[wet, 0] += hump[0][wet, 0]
    st = (ze, cold[1], cold[2], cold[3])
    prob = adjoint.twin_problem(case, nsteps=nst, ng=20, seed=9, state0=st, perturb=0.5)
    prob.data = np.zeros_like(prob.data)
    J, gn, gdp = prob.gradient(case.manning, case.dp)
    o = Oracle(case, threads=8)
    o.set_state(*st)
    o.limiter_record(nst)
    o.step(nst)
    records = o.limiter_records()
    kinds = np.bincount(records[2] & 3, minlength=4)
    sides = np.bincount(records[3], minlength=5)
    print(f"p={p}: WETDRY branches {kinds.tolist()}, dry-side cases {sides.tolist()}", flush=True)
    assert kinds[2] + kinds[3] > 0
    loss = frozen_oracle_loss(case, prob, 20, 9, records)
    assert loss(case.manning, case.dp) == pytest.approx(J, rel=1e-9, abs=0.0)
    for vn, vd in directions(case, 3, seed=14):
        ad = float(gn @ vn + gdp @ vd)
        fd = fd4(lambda t: loss(case.manning + t * vn, case.dp + t * vd), 1e-6)
        rel = abs(fd - ad) / abs(ad)
        print(f"p={p}: adjoint {ad:.12e} frozen-branch oracle fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= TOL
    prob.solver.close()
