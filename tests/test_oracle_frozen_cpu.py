"""The oracle's frozen-branch map (the map the discrete adjoint differentiates): SLOPELIMITER5,
WETDRY and the internal edges' dry-side treatment with every branch recorded from a run and
replayed.

* Replayed at the recording point, the frozen map gives the free run's states and wet/dry flags
  bit for bit, on a basin where elements dry and flood during the run, at p=1 and p=2.
* WETDRY's flat and redistribution branches occur in that run (and at p=1 the check alone), so
  the replay is exercised beyond the identity.
* Replayed at perturbed parameters the frozen map moves smoothly: its second difference in a
  Manning-and-depth direction scales as the step squared at steps of 1e-5 (a switched branch would
  make it scale linearly).  The frozen map still has the kinks the adjoint linearises without
  recording them (the LLF bound, the friction cap), which steps of 1e-4 and more cross on this
  basin.
* From the cold start of a basin with a shelf (dry land along the coast, the water at rest), the
  fluxes across the shoreline are zero, so no edge takes a dry-side case (reflection or
  zero-gravity flux); perturbed parameters make them nonzero and switch thousands of edge points
  to those cases, a jump of order g*H0^2/2 in the momentum flux.  With the cases replayed the
  frozen map stays smooth there.
"""
import copy

import numpy as np
import pytest

from dgswem_amd import adjoint, mesh
from oracle.oracle import Oracle


def drying_basin(p):
    c = mesh.structured_basin(nx=12, ny=6, dx=1000.0, dy=1000.0, depth=10.0, p=p, slopeflag=5,
                              manning=None, dt=2.0 if p == 2 else 5.0, amp=0.0, nolifa=2)
    nid = lambda i, j: i * 7 + j + 1
    ring = [nid(i, 0) for i in range(13)] + [nid(12, j) for j in range(1, 7)] + \
           [nid(i, 6) for i in range(11, -1, -1)] + [nid(0, j) for j in range(5, -1, -1)]
    c.open_segs = []
    c.land_segs = [(0, np.array(ring, dtype=np.int32))]
    c.params.nbfr = 0
    c.params.nramp = 0
    c.dp = 2.0 - 0.0004 * c.x                      # 2 m deep in the west to -2.8 m in the east
    return c


def wetdry_state(c, amplitude):
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
    return ze, qx, qy, np.where(dry, 0, 1).astype(wd.dtype)


STEPS = {1: 120, 2: 240}


@pytest.mark.parametrize("p", [1, 2])
def test_replay_at_the_recording_point_is_the_free_run(p):
    c = drying_basin(p)
    st = wetdry_state(c, 2.0)
    free = Oracle(c, threads=4)
    free.set_state(*st)
    free.step(STEPS[p])
    rec = Oracle(c, threads=4)
    rec.set_state(*st)
    rec.limiter_record(STEPS[p])
    rec.step(STEPS[p])
    records = rec.limiter_records()
    kinds = np.bincount(records[2] & 3, minlength=4)
    print(f"p={p}: WETDRY branches identity/check/flat/redistribution {kinds.tolist()}", flush=True)
    assert kinds[2] > 0 and kinds[3] > 0 and (p == 2 or kinds[1] > 0)
    flags = ((records[2] >> 13) & 1).reshape(-1, c.ne)
    assert (flags != st[3][None, :]).any()          # some element wets or dries during the run
    rep = Oracle(c, threads=4)
    rep.set_state(*st)
    rep.limiter_replay(records)
    rep.step(STEPS[p])
    for a, b, r in zip(free.state(), rep.state(), rec.state()):
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a, r)


def test_frozen_map_is_smooth_at_perturbed_parameters():
    c = drying_basin(1)
    st = wetdry_state(c, 2.0)
    o = Oracle(c, threads=4)
    o.set_state(*st)
    o.limiter_record(60)
    o.step(60)
    records = o.limiter_records()
    rng = np.random.default_rng(3)
    vn = rng.standard_normal(c.np_) * 0.02
    vd = rng.standard_normal(c.np_) * 0.05

    def run(t):
        cc = copy.copy(c)
        cc.manning = (c.manning if c.manning is not None else np.full(c.np_, 0.025)) + t * vn
        cc.dp = c.dp + t * vd
        oc = Oracle(cc, threads=4)
        oc.set_state(*st)
        oc.limiter_replay(records)
        oc.step(60)
        return oc.state()[0]

    u0 = run(0.0)
    second = []
    for h in (2e-5, 1e-5):
        second.append(np.linalg.norm(run(h) - 2.0 * u0 + run(-h)))
    ratio = second[0] / second[1]
    print(f"second differences {second}, ratio {ratio:.2f} (4 for a smooth map)", flush=True)
    assert 3.0 < ratio < 5.0


def test_frozen_map_is_smooth_from_a_cold_start():
    c = mesh.delaunay_basin(n_elem=3000, seed=5, lx=6.0e4, ly=3.6e4, depth_range=(2.0, 20.0), shelf=True,
                            p=1, slopeflag=5, wetdry=True, nsteps=60, cf=1e-4)
    st = Oracle(c, threads=4).state()
    assert (st[3] == 0).any()                       # the cold start's dry check dried the coast
    rng = np.random.default_rng(3)
    vn = rng.standard_normal(c.np_) * c.manning
    vd = rng.standard_normal(c.np_) * c.dp

    def run(t, records=None):
        cc = copy.copy(c)
        cc.manning = c.manning + t * vn
        cc.dp = c.dp + t * vd
        oc = Oracle(cc, threads=4)
        oc.set_state(*st)
        if records is None:
            oc.limiter_record(60)
        else:
            oc.limiter_replay(records)
        oc.step(60)
        return oc

    o = run(0.0)
    records = o.limiter_records()
    assert not records[3].any()                     # no dry-side case at the recording point
    sides = np.bincount(run(1e-5).limiter_records()[3], minlength=5)
    print(f"dry-side cases of the free run at step 1e-5: {sides.tolist()}", flush=True)
    assert sides[1:].sum() > 1000                   # ... but many once the parameters move
    u0 = run(0.0, records).state()[0]
    np.testing.assert_array_equal(u0, o.state()[0])
    second = [np.linalg.norm(run(h, records).state()[0] - 2.0 * u0 + run(-h, records).state()[0])
              for h in (2e-5, 1e-5)]
    ratio = second[0] / second[1]
    print(f"second differences {second}, ratio {ratio:.2f} (4 for a smooth map)", flush=True)
    assert 3.0 < ratio < 5.0
