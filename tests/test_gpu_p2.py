"""The p=2 path on the GPU (six modes per field, six area points, four Lobatto edge points) against
the oracle, and the batched p=2 storm ensemble (the 32 x 100k p=2 workload in miniature).

The reference's own p=2 run diverges (its p=2 edge tables are p=1-only), so the p=2 oracle is
pinned by invariants instead (lake at rest, mass, mirror symmetry, convergence order against the
closed-form standing wave: tests/test_oracle_invariants.py).  Tolerances: bit for bit where the
arithmetic has no transcendental function (sqrt and division round correctly on both sides);
relative L2 <= 1e-10 with tides (device cos) or Holland forcing (device sin/cos/exp/pow).
"""
import numpy as np
import pytest

from dgswem_amd import adjoint, gpu, mesh
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def tidal_p2(nsteps=200):
    return mesh.delaunay_basin(n_elem=3000, seed=41, lx=2.0e5, ly=1.2e5, depth_range=(5.0, 200.0),
                               shelf=False, wetdry=False, slopeflag=0, p=2, cfl=0.15, nsteps=nsteps)


def closed_p2(nsteps=200):
    """Closed 24 x 12 structured basin, p=2, sloping bottom, uniform quadratic friction (no
    Manning), no tides: no transcendental function anywhere on the path."""
    nx, ny = 24, 12
    c = mesh.structured_basin(nx=nx, ny=ny, dx=1000.0, dy=1000.0, depth=10.0, p=2, slopeflag=0,
                              manning=None, dt=2.0, amp=0.0)
    nid = lambda i, j: i * (ny + 1) + j + 1
    ring = [nid(i, 0) for i in range(nx + 1)] + [nid(nx, j) for j in range(1, ny + 1)] + \
           [nid(i, ny) for i in range(nx - 1, -1, -1)] + [nid(0, j) for j in range(ny - 1, -1, -1)]
    c.open_segs = []
    c.land_segs = [(0, np.array(ring, dtype=np.int32))]
    c.params.nbfr = 0
    c.params.nramp = 0
    c.dp = 8.0 + 0.0002 * c.x + np.sin(c.y / 2500.0)
    c.params.nt = nsteps
    return c


def run_both(case, steps, state0=None):
    g = gpu.GpuSolver(case)
    o = Oracle(case, threads=8)
    if state0 is not None:
        g.set_state(*state0)
        o.set_state(*state0)
    g.step(steps)
    o.step(steps)
    a, b = g.state(), o.state()
    g.close()
    return a, b


def test_p2_tidal_basin_matches_oracle():
    case = tidal_p2()
    print(f"p=2 tidal basin: {case.ne} elements, dt {case.params.dtdp:.4f} s", flush=True)
    a, b = run_both(case, 200, adjoint.hump_state(case, amplitude=0.3))
    np.testing.assert_array_equal(a[3], b[3])
    for x, y, name in zip(a[:3], b[:3], ("ze", "qx", "qy")):
        r = rel_l2(x, y)
        print(f"{name} rel L2 {r:.3e}", flush=True)
        assert np.isfinite(x).all() and r <= TOL


def test_p2_closed_basin_bitwise():
    case = closed_p2()
    ze, qx, qy, wd = adjoint.hump_state(case, amplitude=0.2, width=0.2)
    a, b = run_both(case, 200, (ze, qx, qy, wd))
    assert np.abs(a[0]).max() > 1e-3          # the hump has moved
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_p2_limiter_wetdry_matches_oracle():
    """The bench configuration at p=2 in small: limiter 5, wetting and drying over a shelf that rises
    onshore, five tidal constituents, Manning n per node."""
    case = mesh.delaunay_basin(n_elem=4000, seed=43, lx=2.0e5, ly=1.2e5, p=2, cfl=0.15, nsteps=300)
    a, b = run_both(case, 300)
    assert (b[3] == 0).sum() > 0                  # some elements are dry
    np.testing.assert_array_equal(a[3], b[3])
    for x, y, name in zip(a[:3], b[:3], ("ze", "qx", "qy")):
        r = rel_l2(x, y)
        print(f"{name} rel L2 {r:.3e}", flush=True)
        assert r <= TOL


def test_p2_block_sizes_bitwise():
    case = tidal_p2(nsteps=100)
    st = adjoint.hump_state(case, amplitude=0.3)
    out = []
    for nt in (64, 128):
        g = gpu.GpuSolver(case)
        g.set_state(*st)
        g.set_block(nt)
        g.step(100)
        out.append(g.state())
        g.close()
    for x, y in zip(*out):
        np.testing.assert_array_equal(x, y)


def storm_members():
    return [mesh.storm_basin(ne, seed=seed, nsteps=100, p=2, slopeflag=0, wetdry=False, shelf=False,
                             cfl=0.15, lx=3.0e5, ly=2.0e5, depth_range=(5.0, 200.0))
            for ne, seed in ((2500, 51), (4000, 52), (1800, 53))]


def test_p2_storm_ensemble_matches_single_runs_and_oracle():
    cases = storm_members()
    steps = 100
    print(f"p=2 storm ensemble of {[c.ne for c in cases]} elements, {steps} steps", flush=True)
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
        print(f"member {i}", flush=True)
        for x, y in zip(p, s):
            np.testing.assert_array_equal(x, y)
    ens.close()
    o = Oracle(cases[0], threads=8)
    o.step(steps)
    ref = o.state()
    assert np.abs(ref[1]).max() > 1e-3            # the storm has forced a flow
    for x, y, name in zip(solo[0][:3], ref[:3], ("ze", "qx", "qy")):
        r = rel_l2(x, y)
        print(f"member 0 vs oracle: {name} rel L2 {r:.3e}", flush=True)
        assert r <= TOL
