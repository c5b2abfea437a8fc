"""Batched ensembles, HIP-graph replay and the block-size autotuner (GPU, through the C ABI).

Batching, graph replay and the block size change only how launches are grouped, never the
arithmetic of an element, so every comparison with the single-run eager path is bit for bit.
The first member is the reference fixture `delaunay4k`, checked against dgswem_serial's FP64 state
to the parity tolerance (relative L2 <= 1e-10, identical wet/dry mask).
"""
import numpy as np
import pytest

from conftest import golden, load_case
from dgswem_amd import gpu, mesh

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def member_cases():
    """Uneven meshes with the switches of `delaunay4k` (limiter 5, wet/dry, Manning) but their own
    sizes, time steps, bathymetry and tidal amplitudes."""
    cases = [load_case("delaunay4k")]
    for ne, seed in ((1500, 11), (6000, 12), (2500, 13)):
        cases.append(mesh.delaunay_basin(n_elem=ne, seed=seed, nsteps=100))
    return cases


def assert_same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_ensemble_matches_single_runs_bitwise():
    cases = member_cases()
    steps = 100
    print(f"ensemble of {[c.ne for c in cases]} elements, {steps} steps", flush=True)
    solo = []
    for c in cases:
        g = gpu.GpuSolver(c)
        g.step(steps)
        solo.append(g.state())
        g.close()
    members = [gpu.GpuSolver(c) for c in cases]
    ens = gpu.Ensemble(members)
    assert ens.nmembers == len(cases)
    np.testing.assert_array_equal(np.diff(ens.eoff), [c.ne for c in cases])
    np.testing.assert_array_equal(np.diff(ens.noff), [c.np_ for c in cases])
    ens.step(steps)
    parts = ens.split(ens.state())
    for i, (p, s) in enumerate(zip(parts, solo)):
        print(f"member {i}", flush=True)
        assert_same(p, s)
    ref = golden(f"delaunay4k_step{steps}")
    ze, qx, qy, wd = parts[0]
    np.testing.assert_array_equal(wd, ref["wdflg"])
    for a, b in ((ze, ref["ze"]), (qx, ref["qx"]), (qy, ref["qy"])):
        assert rel_l2(a, b) <= 1e-10
    # member handles are untouched by the ensemble
    assert_same(members[1].state(), gpu.GpuSolver(cases[1]).state())
    ens.close()
    for m in members:
        m.close()


def test_ensemble_nodal_and_mass_split_by_member():
    cases = member_cases()[:3]
    members = [gpu.GpuSolver(c) for c in cases]
    for m in members:
        m.step(20)
    ens = gpu.Ensemble(members)
    assert ens.it == 20
    ens.step(30)
    for m in members:
        m.step(30)
    eta, uu, vv, nc = ens.nodal()
    for i, part in enumerate(ens.split((eta, uu, vv, nc), nodes=True)):
        assert_same(part, members[i].nodal())
    total = sum(m.mass() for m in members)
    assert abs(ens.mass() - total) <= 1e-12 * abs(total)
    ens.close()
    for m in members:
        m.close()


def test_ensemble_with_graph_matches_eager():
    cases = member_cases()
    a = gpu.Ensemble([gpu.GpuSolver(c) for c in cases])
    b = gpu.Ensemble([gpu.GpuSolver(c) for c in cases])
    b.use_graph(True)
    for n in (7, 50, 1, 45):
        print(f"steps {n}", flush=True)
        a.step(n)
        b.step(n)
        assert_same(a.state(), b.state())
    a.close()
    b.close()


def test_graph_replay_matches_eager_single_run():
    case = load_case("wetdry")   # SSP(5,2): 6-step slot-rotation period, wet/dry fronts
    a, b = gpu.GpuSolver(case), gpu.GpuSolver(case)
    b.use_graph(True)
    for n in (13, 6, 40):
        a.step(n)
        b.step(n)
        assert_same(a.state(), b.state())
    b.use_graph(False)
    a.step(5)
    b.step(5)
    assert_same(a.state(), b.state())
    a.close()
    b.close()


def test_autotune_and_block_sizes_do_not_change_results():
    case = load_case("delaunay4k")
    g = gpu.GpuSolver(case)
    g.step(10)
    s0 = g.state()
    best = g.autotune()
    print(f"autotuned block {best}", flush=True)
    assert best in (64, 128, 256)
    assert_same(g.state(), s0)     # autotune touches only scratch stage buffers
    g.step(25)
    ref = g.state()
    for block in (64, 128, 256):
        h = gpu.GpuSolver(case)
        h.step(10)
        h.set_block(block)
        h.step(25)
        assert_same(h.state(), ref)
        h.close()
    g.close()


def test_ensemble_errors():
    c1 = load_case("c1")
    wd = load_case("wetdry")
    a, b = gpu.GpuSolver(c1), gpu.GpuSolver(wd)
    with pytest.raises(gpu.GpuError, match="share"):
        gpu.Ensemble([a, b])
    a2 = gpu.GpuSolver(c1)
    a2.step(1)
    with pytest.raises(gpu.GpuError, match="same time step"):
        gpu.Ensemble([a, a2])
    ens = gpu.Ensemble([a, gpu.GpuSolver(c1)])
    p = c1.params
    L = gpu.load_library()
    emo, efa = gpu.emo_dg(c1)
    rc = L.dg_gpu_set_tide(ens.h, p.nbfr, gpu._f(p.amig), gpu._f(p.ff), gpu._f(p.face),
                           gpu._f(p.per), gpu._f(emo), gpu._f(efa))
    assert rc != 0 and b"dg_ens_create" in L.dg_gpu_last_error()
    with pytest.raises(gpu.GpuError, match="block size"):
        ens.set_block(96)
    with pytest.raises(gpu.GpuError, match="block size"):
        ens.set_block(192)
    ens.close()
    for h in (a, b, a2):
        h.close()


def test_storm_ensemble_matches_single_runs_bitwise():
    """Members under their own Holland hurricanes (NWS=8, spherical, variable Coriolis): the node
    forcing kernel runs once per step over all members' nodes."""
    cases = [load_case("storm3k")]
    for ne, seed in ((2000, 21), (5000, 22)):
        cases.append(mesh.storm_basin(ne, seed=seed, nsteps=60, lx=3.0e5, ly=2.0e5,
                                      depth_range=(5.0, 200.0)))
    steps = 60
    solo = []
    for c in cases:
        g = gpu.GpuSolver(c)
        g.step(steps)
        solo.append(g.state())
        g.close()
    ens = gpu.Ensemble([gpu.GpuSolver(c) for c in cases])
    ens.use_graph(True)
    ens.step(steps)
    for i, (p, s) in enumerate(zip(ens.split(ens.state()), solo)):
        print(f"storm member {i}", flush=True)
        assert_same(p, s)
    ens.close()


def test_storm_ensemble_rejects_mixed_forcing():
    a = gpu.GpuSolver(load_case("storm3k"))
    b = gpu.GpuSolver(mesh.storm_basin(2000, seed=21, nsteps=10, lx=3.0e5, ly=2.0e5,
                                       depth_range=(5.0, 200.0)))
    c = mesh.delaunay_basin(n_elem=2000, seed=21, nsteps=10, lx=3.0e5, ly=2.0e5,
                            depth_range=(5.0, 200.0), ics=2)
    nc = gpu.GpuSolver(c)
    with pytest.raises(gpu.GpuError, match="NWS=8"):
        gpu.Ensemble([a, nc])
    gpu.Ensemble([a, b]).close()
    for h in (a, b, nc):
        h.close()


@pytest.mark.parametrize("name,steps", [("delaunay4k", 40), ("delaunay4k_nolim", 40), ("storm3k", 40),
                                        ("c1var", 40), ("quarter_annular", 30), ("mass_conservation", 40)])
def test_every_launch_shape_is_bitwise_identical(name, steps):
    """Every block size across the kernel's variants: limiter + wet/dry, wet/dry fused into the
    stage (no limiter), NWS=8 with spherical coordinates, a flux boundary, SSP(5,2) with kept stage
    RHS, rain."""
    case = load_case(name)
    shapes = (64, 128, 256) if case.dg.ph == 1 else (64, 128)
    ref = None
    for b in shapes:
        print(f"{name}: block {b}", flush=True)
        g = gpu.GpuSolver(case)
        g.set_block(b)
        g.step(steps)
        st = g.state()
        g.close()
        if ref is None:
            ref = st
        else:
            assert_same(st, ref)
