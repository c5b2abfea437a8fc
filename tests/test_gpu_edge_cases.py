"""Edge cases of the GPU step's tiling and boundary handling, against the oracle bit for bit.

Closed basins (land on every side, no tide), uniform quadratic friction (no Manning cube root):
no transcendental function is on the path, so the GPU state must equal the oracle's exactly.

* two elements sharing one internal edge (one tile, a one-element halo, one edge task);
* 66 elements: one full 64-element tile and a two-element tail tile (lanes past the last element);
* a basin whose every node lies above the datum with wetting/drying on: every element dry for
  the whole run (no internal flux at all);
* a basin that dries and floods: land rising across the basin under a moving hump;
* the same kind of basin with its elements and nodes numbered at random: every tile's neighbours
  lie all over the mesh (the stage kernel's halos are as large as they get, and a 256-element
  limiter tile touches up to 692 nodes and 2220 halo elements: 90 KB of LDS);
each at p=1 and p=2 and at every stage block size.
"""
import copy

import numpy as np
import pytest

from dgswem_amd import adjoint, gpu, mesh
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def closed(nx, ny, p, depth=10.0, nolifa=1, slopeflag=0):
    c = mesh.structured_basin(nx=nx, ny=ny, dx=1000.0, dy=1000.0, depth=depth, p=p, slopeflag=slopeflag,
                              manning=None, dt=2.0 if p == 2 else 5.0, amp=0.0, nolifa=nolifa)
    nid = lambda i, j: i * (ny + 1) + j + 1
    ring = [nid(i, 0) for i in range(nx + 1)] + [nid(nx, j) for j in range(1, ny + 1)] + \
           [nid(i, ny) for i in range(nx - 1, -1, -1)] + [nid(0, j) for j in range(ny - 1, -1, -1)]
    c.open_segs = []
    c.land_segs = [(0, np.array(ring, dtype=np.int32))]
    c.params.nbfr = 0
    c.params.nramp = 0
    return c


def wetdry_state(c, amplitude):
    """A hump where the water is deep enough, and elsewhere WETDRY's dry state: the free surface
    parallel to the bottom, depth H0/2, no flow, dry flag."""
    ze, qx, qy, wd = adjoint.hump_state(c, amplitude=amplitude, width=0.3)
    d = c.dp[c.nm - 1]
    hb = np.stack([(d[:, 0] + d[:, 1] + d[:, 2]) / 3.0, -(d[:, 0] + d[:, 1]) / 6.0 + d[:, 2] / 3.0,
                   -(0.5 * d[:, 0]) + 0.5 * d[:, 1]], axis=1)
    dry = ze[:, 0] + d.min(axis=1) <= 10.0 * c.params.h0
    ze[dry] = 0.0
    ze[dry, 0] = 0.5 * c.params.h0 - hb[dry, 0]
    ze[dry, 1] = -hb[dry, 1]
    ze[dry, 2] = -hb[dry, 2]
    qx[dry] = 0.0
    qy[dry] = 0.0
    wd = np.where(dry, 0, 1).astype(wd.dtype)
    return ze, qx, qy, wd


def compare(case, steps, state0, blocks):
    o = Oracle(case, threads=4)
    o.set_state(*state0)
    o.step(steps)
    ref = o.state()
    for nt in blocks:
        g = gpu.GpuSolver(case)
        g.set_state(*state0)
        g.set_block(nt)
        g.step(steps)
        got = g.state()
        g.close()
        print(f"  block {nt}", flush=True)
        for x, y in zip(got, ref):
            np.testing.assert_array_equal(x, y)
    return ref


BLOCKS = {1: (64, 128, 256), 2: (64, 128)}


@pytest.mark.parametrize("p", [1, 2])
def test_two_elements(p):
    c = closed(1, 1, p)
    print(f"two elements p={p}", flush=True)
    ze, qx, qy, wd = adjoint.rest_state(c)
    ze[0, 0], ze[1, 0] = 0.2, -0.1
    ref = compare(c, 50, (ze, qx, qy, wd), BLOCKS[p])
    assert np.abs(ref[1]).max() > 0.0


@pytest.mark.parametrize("p", [1, 2])
def test_partial_last_tile(p):
    c = closed(3, 11, p, slopeflag=5)
    assert c.ne == 66
    print(f"66 elements p={p}", flush=True)
    compare(c, 40, adjoint.hump_state(c, amplitude=0.3, width=0.3), BLOCKS[p])


@pytest.mark.parametrize("p", [1, 2])
def test_all_dry(p):
    c = closed(6, 4, p, depth=-1.0, nolifa=2, slopeflag=5)
    print(f"all dry p={p}", flush=True)
    st = wetdry_state(c, 0.0)
    assert (st[3] == 0).all()
    ref = compare(c, 30, st, BLOCKS[p])
    assert (ref[3] == 0).all()


@pytest.mark.parametrize("p", [1, 2])
def test_drying_and_flooding(p):
    c = closed(12, 6, p, nolifa=2, slopeflag=5)
    c.dp = 2.0 - 0.0004 * c.x                      # 2 m deep in the west to -2.8 m in the east
    print(f"drying and flooding p={p}", flush=True)
    st = wetdry_state(c, 2.0)
    # the oracle in chunks: some elements wet or dry during the run
    o = Oracle(c, threads=4)
    o.set_state(*st)
    prev, changes = st[3].copy(), 0
    for _ in range(8):
        o.step(25)
        w = o.state()[3]
        changes += int((w != prev).sum())
        prev = w.copy()
    assert (st[3] == 0).any() and changes > 0
    ref = compare(c, 200, st, BLOCKS[p])
    np.testing.assert_array_equal(ref[3], prev)
    assert np.isfinite(ref[0]).all()


def renumbered(c, seed):
    """The case with its elements and nodes in a seeded random order."""
    rng = np.random.default_rng(seed)
    r = copy.deepcopy(c)
    pn = rng.permutation(c.np_)                      # new node k is old node pn[k]
    inv = np.empty_like(pn)
    inv[pn] = np.arange(c.np_)
    for f in ("x", "y", "slam", "sfea", "dp", "sfac", "corif"):
        setattr(r, f, getattr(c, f)[pn])
    if c.manning is not None:
        r.manning = c.manning[pn]
    r.nm = (inv[c.nm - 1] + 1).astype(np.int32)[rng.permutation(c.ne)]
    r.open_segs = [(inv[s_ - 1] + 1).astype(np.int32) for s_ in c.open_segs]
    r.land_segs = [(t, (inv[s_ - 1] + 1).astype(np.int32)) for t, s_ in c.land_segs]
    return r


@pytest.mark.parametrize("p", [1, 2])
def test_random_numbering(p):
    c = renumbered(closed(60, 40, p, nolifa=2, slopeflag=5), seed=5)
    c.dp = 4.0 - 0.0001 * c.x                      # 4 m deep in the west to -2 m in the east
    print(f"random numbering p={p}: {c.ne} elements", flush=True)
    ref = compare(c, 40, wetdry_state(c, 1.0), BLOCKS[p])
    assert np.isfinite(ref[0]).all() and (ref[3] == 0).any()
