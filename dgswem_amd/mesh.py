"""Seeded synthetic meshes and run directories for the DG hydro step.

* `structured_basin` -- rectangular basin of nx x ny squares, each split into two triangles, with one
  open (tidal) side and land elsewhere: the small plain-method case and its variants.
* `delaunay_basin` -- seeded Delaunay mesh of jittered random points refined toward one coast, with a
  smooth Gaussian-random-field bathymetry and a shelf that rises above datum onshore (wet/dry), a
  tidal open boundary with several constituents, per-node Manning n, and a CFL-limited time step.
  Nodes and elements are ordered along a Hilbert curve so that neighbouring elements sit close in
  memory, which is what the element-major device layout relies on.
* `write_run_dir` -- writes fort.14/fort.15/fort.13/fort.dg for a Case so that the reference's
  serial solver can run the very same input.

Everything is deterministic for a given seed.
"""
from __future__ import annotations

import math
import os
from typing import Optional

import numpy as np

from .fortio import DEG2RAD, PI, R_EARTH, Case, DGOptions, Params

# Tidal constituent angular frequencies (rad/s): M2, S2, N2, K1, O1
CONSTITUENTS = [("M2", 1.405189e-4), ("S2", 1.454441e-4), ("N2", 1.378797e-4),
                ("K1", 7.292117e-5), ("O1", 6.759774e-5)]


def _params(**kw) -> Params:
    p = Params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def structured_basin(nx: int = 40, ny: int = 20, dx: float = 2500.0, dy: float = 2500.0,
                     depth: float = 10.0, amp: float = 0.5, manning: Optional[float] = 0.025,
                     dt: float = 10.0, nsteps: int = 8640, p: int = 1, rk_stage: int = 2,
                     slopeflag: int = 0, nolifa: int = 1, cori: float = 0.0, nolibf: int = 1,
                     flow_side: bool = False, flow_amp: float = 0.0) -> Case:
    """Rectangular basin; node id = i*(ny+1)+j+1 at (i*dx, j*dy); squares split (a,b,d),(b,c,d).

    Open boundary on x=0 listed from (0,ny) down to (0,0) and forced by an M2 tide of amplitude
    `amp` with zero phase; the other three sides are one no-normal-flow segment.  With
    `flow_side` the x=nx*dx side becomes a flux-specified segment (type 2) with an M2 normal flux.
    """
    n_nodes = (nx + 1) * (ny + 1)
    x = np.empty(n_nodes); y = np.empty(n_nodes)
    for i in range(nx + 1):
        for j in range(ny + 1):
            k = i * (ny + 1) + j
            x[k] = dx * i; y[k] = dy * j
    nid = lambda i, j: i * (ny + 1) + j + 1
    nm = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            nm.append((a, b, d)); nm.append((b, c, d))
    nm = np.array(nm, dtype=np.int32)
    open_seg = np.array([nid(0, j) for j in range(ny, -1, -1)], dtype=np.int32)
    if flow_side:
        land1 = np.array([nid(i, 0) for i in range(0, nx + 1)], dtype=np.int32)
        flow = np.array([nid(nx, j) for j in range(0, ny + 1)], dtype=np.int32)
        land2 = np.array([nid(i, ny) for i in range(nx, -1, -1)], dtype=np.int32)
        land_segs = [(0, land1), (2, flow), (0, land2)]
    else:
        land = [nid(i, 0) for i in range(0, nx + 1)] + [nid(nx, j) for j in range(1, ny + 1)] + \
               [nid(i, ny) for i in range(nx - 1, -1, -1)]
        land_segs = [(0, np.array(land, dtype=np.int32))]
    w = CONSTITUENTS[0][1]
    prm = _params(ics=1, nolibf=nolibf, nolifa=nolifa, nolica=1, nolicat=1, nwp=1 if manning else 0,
                  ncor=0, ntip=0, nws=0, nramp=1, g=9.81, tau0=0.005, dtdp=dt, statim=0.0,
                  reftim=0.0, rnday=nsteps * dt / 86400.0, dramp=0.25, h0=0.01,
                  cf=0.0025, tau=0.0025 if nolibf == 0 else 0.0,
                  cori=cori, nt=nsteps, nbfr=1,
                  amig=np.array([w]), ff=np.array([1.0]), face=np.array([0.0]),
                  per=np.array([2.0 * PI / w]), emo=np.full((1, ny + 1), amp),
                  efa=np.zeros((1, ny + 1)), noutge=1, toutsge=0.0, toutfge=nsteps * dt / 86400.0,
                  nspoolge=nsteps)
    if flow_side:
        prm.nffr = 1
        prm.famig = np.array([w]); prm.fff = np.array([1.0]); prm.fface = np.array([0.0])
        prm.fper = np.array([2.0 * PI / w])
        prm.qnam = np.full((1, ny + 1), flow_amp); prm.qnph = np.zeros((1, ny + 1))
    dg = DGOptions(pl=p, ph=p, fluxtype=2, rk_stage=rk_stage, rk_order=2, slopeflag=slopeflag)
    dpn = np.full(n_nodes, depth)
    if nolifa in (0, 1):
        dpn = np.where(dpn < prm.h0, prm.h0, dpn)
    return Case(title="basin", params=prm, dg=dg, x=x, y=y,
                slam=x / (6378206.4 * math.cos(0.0)), sfea=y / 6378206.4, dp=dpn,
                sfac=np.ones(n_nodes), corif=np.full(n_nodes, cori), nm=nm,
                open_segs=[open_seg], land_segs=land_segs,
                manning=np.full(n_nodes, manning) if manning else None)


# ----------------------------------------------------------------------------------------------
# Delaunay basin
# ----------------------------------------------------------------------------------------------
def _hilbert_index(xi: np.ndarray, yi: np.ndarray, order: int) -> np.ndarray:
    """Hilbert-curve index of integer grid coordinates in [0, 2**order)."""
    x = xi.astype(np.int64).copy(); y = yi.astype(np.int64).copy()
    d = np.zeros_like(x)
    s = 1 << (order - 1)
    while s > 0:
        rx = ((x & s) > 0).astype(np.int64)
        ry = ((y & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate
        m = ry == 0
        flip = m & (rx == 1)
        x = np.where(flip, s - 1 - x, x); y = np.where(flip, s - 1 - y, y)
        x2 = np.where(m, y, x); y2 = np.where(m, x, y)
        x, y = x2, y2
        s >>= 1
    return d


def hilbert_order(px: np.ndarray, py: np.ndarray, order: int = 16) -> np.ndarray:
    n = (1 << order) - 1
    x0, x1 = px.min(), px.max(); y0, y1 = py.min(), py.max()
    span = max(x1 - x0, y1 - y0, 1e-30)
    xi = np.clip(((px - x0) / span * n).astype(np.int64), 0, n)
    yi = np.clip(((py - y0) / span * n).astype(np.int64), 0, n)
    return np.argsort(_hilbert_index(xi, yi, order), kind="stable")


def _grf(rng, nx: int, ny: int, corr: float) -> np.ndarray:
    """Smooth Gaussian random field on an nx x ny grid (Gaussian spectral filter), unit variance."""
    w = rng.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx)[None, :]; ky = np.fft.fftfreq(ny)[:, None]
    filt = np.exp(-0.5 * (kx ** 2 + ky ** 2) * (2 * PI * corr) ** 2)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    f -= f.mean()
    return f / (f.std() + 1e-30)


def _bilinear(grid: np.ndarray, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    ny, nx = grid.shape
    fx = np.clip(u * (nx - 1), 0, nx - 1 - 1e-9); fy = np.clip(v * (ny - 1), 0, ny - 1 - 1e-9)
    i = fx.astype(np.int64); j = fy.astype(np.int64); a = fx - i; b = fy - j
    return ((1 - a) * (1 - b) * grid[j, i] + a * (1 - b) * grid[j, i + 1]
            + (1 - a) * b * grid[j + 1, i] + a * b * grid[j + 1, i + 1])


def delaunay_basin(n_elem: int = 1_000_000, seed: int = 0, lx: float = 1.0e6, ly: float = 6.0e5,
                   depth_range=(5.0, 4000.0), shelf: bool = True, shelf_frac: float = 0.06,
                   land_height: float = 2.0, n_const: int = 5, p: int = 1, cfl: float = 0.3,
                   slopeflag: int = 5, wetdry: bool = True, nsteps: int = 2000,
                   manning_range=(0.02, 0.03), refine: float = 3.0, ics: int = 1,
                   lon_c: float = -75.03, lat_c: float = 27.97, storm=None,
                   cf: float = 0.0025) -> Case:
    """Seeded Delaunay basin of about `n_elem` triangles (see module docstring).

    The open boundary is x=0 (west); the coast x=lx (east) is land with a shelf that rises from the
    random-field depth to `land_height` metres above datum (DP=-land_height) over the last
    `shelf_frac` of the domain; north and south sides are land.  Point density grows by a factor
    `1+refine` toward the coast.  The time step is `cfl` times the smallest element inradius over
    the gravity wave speed of its deepest node.  `cf` is fort.15's CF, the floor of the Manning
    friction coefficient g n^2 / H^(1/3).

    With ics=2 the basin is placed on the sphere, centred at (lon_c, lat_c) degrees, which is also
    the CPP projection centre; nodes are stored as fort.14 lon/lat and projected exactly as the
    reference's reader does, and the Coriolis parameter varies with latitude (NCOR=1).  `storm`
    (a holland.HollandStorm) switches on NWS=8 Holland vortex forcing.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    n_nodes = n_elem // 2 + 1
    # Jittered hexagonal columns whose spacing h(x) = h0/sqrt(1+refine*(x/lx)^2) shrinks smoothly
    # toward the coast; the first and last point of every column sit on y=0 / y=ly and the first
    # and last column on x=0 / x=lx, so the boundary is point-exact.
    dens = lambda xx: 1.0 + refine * (xx / lx) ** 2

    def columns(h0):
        xs = [0.0]
        while True:
            hx = h0 / math.sqrt(dens(xs[-1])) * math.sqrt(3.0) / 2.0
            if xs[-1] + hx >= lx - 0.5 * hx:
                break
            xs.append(xs[-1] + hx)
        xs = np.array(xs + [lx])
        # stretch the columns so the last gap is regular
        xs = xs * (lx / xs[-1])
        return xs

    def count(h0):
        xs = columns(h0)
        return sum(max(2, int(round(ly / (h0 / math.sqrt(dens(xx))))) + 1) for xx in xs)

    h0 = math.sqrt(lx * ly / n_nodes * 2.0 / math.sqrt(3.0) / (1.0 + refine / 3.0))
    for _ in range(4):
        h0 *= math.sqrt(count(h0) / n_nodes)
    xs = columns(h0)
    pts, west, east, south, north = [], [], [], [], []
    for ci, xx in enumerate(xs):
        hy = h0 / math.sqrt(dens(xx))
        ny_ = max(2, int(round(ly / hy))) + 1
        yy = np.linspace(0.0, ly, ny_)
        if ci % 2 == 1 and ny_ > 2:
            yy[1:-1] = yy[1:-1] + 0.5 * (ly / (ny_ - 1)) * np.where(np.arange(1, ny_ - 1) < ny_ - 2, 1.0, 0.0)
        xcol = np.full(ny_, xx)
        jit = 0.12 * (ly / (ny_ - 1))
        if 0 < ci < len(xs) - 1:
            xcol[1:-1] += rng.uniform(-jit, jit, ny_ - 2)
        yy[1:-1] += rng.uniform(-jit, jit, ny_ - 2)
        base = sum(len(q) for q in pts)
        pts.append(np.stack([xcol, yy], axis=1))
        south.append(base); north.append(base + ny_ - 1)
        if ci == 0:
            west = list(range(base, base + ny_))
        if ci == len(xs) - 1:
            east = list(range(base, base + ny_))
    allp = np.concatenate(pts)
    # boundary node strings (pre-permutation indices)
    s_list = south                                  # (0,0) -> (lx,0)
    e_list = east                                   # (lx,0) -> (lx,ly)
    n_list = north[::-1]                            # (lx,ly) -> (0,ly)
    w_list = west[::-1]                             # (0,ly) -> (0,0)
    # Hilbert order for locality
    perm = hilbert_order(allp[:, 0], allp[:, 1])
    inv = np.empty_like(perm); inv[perm] = np.arange(len(perm))
    allp = allp[perm]
    tri = Delaunay(allp).simplices.astype(np.int64)
    X, Y = allp[:, 0], allp[:, 1]
    a = (X[tri[:, 0]] - X[tri[:, 2]]) * (Y[tri[:, 1]] - Y[tri[:, 2]]) + \
        (X[tri[:, 2]] - X[tri[:, 1]]) * (Y[tri[:, 0]] - Y[tri[:, 2]])
    neg = a < 0
    tri[neg] = tri[neg][:, [0, 2, 1]]
    a = np.abs(a)
    tri = tri[a > 1e-9 * np.median(a)]
    cx = X[tri].mean(axis=1); cy = Y[tri].mean(axis=1)
    tri = tri[hilbert_order(cx, cy)]
    nm = (tri + 1).astype(np.int32)
    ids = lambda lst: (inv[np.array(lst, dtype=np.int64)] + 1).astype(np.int32)
    open_seg = ids(w_list)                                              # (0,ly) -> (0,0)
    land_seg = np.concatenate([ids(s_list), ids(e_list[1:]), ids(n_list[1:]), ids(w_list[:1])])
    # bathymetry: GRF mapped to the depth range, then the shelf toward the east coast
    g = _grf(rng, 512, int(512 * ly / lx), corr=0.04 * 512)
    f = _bilinear(g, X / lx, Y / ly)
    f = (f - f.min()) / (f.max() - f.min())
    dp = depth_range[0] + f * (depth_range[1] - depth_range[0])
    if shelf:
        xs = lx * (1.0 - shelf_frac)
        s = np.clip((X - xs) / (lx - xs), 0.0, 1.0)
        dp = dp * (1.0 - s) + (-land_height) * s
    manning = rng.uniform(manning_range[0], manning_range[1], size=len(X))
    # tides: n_const constituents, random amplitudes U(0.1,1) m, phase varying along the boundary
    nbn = len(open_seg)
    const = CONSTITUENTS[:n_const]
    amps = rng.uniform(0.1, 1.0, size=n_const)
    ph0 = rng.uniform(0.0, 360.0, size=n_const)
    emo = np.tile(amps[:, None], (1, nbn))
    efa = (ph0[:, None] + np.linspace(0.0, 20.0, nbn)[None, :]) * DEG2RAD
    # time step from the CFL number, element inradius and the deepest node's wave speed
    x1, x2, x3 = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
    y1, y2, y3 = Y[tri[:, 0]], Y[tri[:, 1]], Y[tri[:, 2]]
    area2 = np.abs((x1 - x3) * (y2 - y3) + (x3 - x2) * (y1 - y3))
    per = np.hypot(x2 - x1, y2 - y1) + np.hypot(x3 - x2, y3 - y2) + np.hypot(x1 - x3, y1 - y3)
    rin = area2 / per
    hmax = np.maximum(np.max(dp[tri], axis=1), 0.5)
    dt = float(cfl * np.min(rin / np.sqrt(9.81 * hmax)))
    prm = _params(ics=ics, nolibf=1, nolifa=2 if wetdry else 1, nolica=1, nolicat=1, nwp=1,
                  ncor=1 if ics == 2 else 0, ntip=0, nws=8 if storm is not None else 0, nramp=1, g=9.81, tau0=0.005, dtdp=dt, statim=0.0, reftim=0.0,
                  rnday=nsteps * dt / 86400.0, dramp=0.5, h0=0.1, cf=cf, cori=0.0,
                  nt=nsteps, nbfr=n_const, amig=np.array([c[1] for c in const]),
                  ff=np.ones(n_const), face=np.zeros(n_const),
                  per=np.array([2.0 * PI / c[1] for c in const]), emo=emo, efa=efa,
                  noutge=1, toutsge=0.0, toutfge=nsteps * dt / 86400.0, nspoolge=nsteps)
    dg = DGOptions(pl=p, ph=p, fluxtype=2, rk_stage=2, rk_order=2, slopeflag=slopeflag)
    if not wetdry:
        dp = np.where(dp < prm.h0, prm.h0, dp)
    if ics == 2:
        from .fortio import coriolis, spherical_coords
        slam0, sfea0 = _exact_radians(lon_c), _exact_radians(lat_c)
        prm.slam0, prm.sfea0 = slam0, sfea0
        c0 = math.cos(sfea0)
        lonlat = np.stack([lon_c + (X - 0.5 * lx) / (R_EARTH * c0) / DEG2RAD,
                           lat_c + (Y - 0.5 * ly) / R_EARTH / DEG2RAD], axis=1)
        slam, sfea, x, y, sfac = spherical_coords(lonlat, slam0, sfea0)
        return Case(title=f"storm{len(nm)}", params=prm, dg=dg, x=x, y=y, slam=slam, sfea=sfea,
                    dp=dp, sfac=sfac, corif=coriolis(sfea), nm=nm, open_segs=[open_seg],
                    land_segs=[(0, land_seg)], manning=manning, storm=storm, lonlat=lonlat)
    if storm is not None:
        raise ValueError("Holland vortex forcing needs spherical coordinates (ics=2)")
    return Case(title=f"delaunay{len(nm)}", params=prm, dg=dg, x=X.copy(), y=Y.copy(),
                slam=X / 6378206.4, sfea=Y / 6378206.4, dp=dp, sfac=np.ones(len(X)),
                corif=np.zeros(len(X)), nm=nm, open_segs=[open_seg], land_segs=[(0, land_seg)],
                manning=manning)


def _exact_radians(deg: float) -> float:
    """DEG2RAD*d for a d near `deg` whose radians survive the fort.15 round trip
    (written as repr(r/DEG2RAD), read back as DEG2RAD*value)."""
    d = float(deg)
    for _ in range(64):
        r = DEG2RAD * d
        if DEG2RAD * float(repr(r / DEG2RAD)) == r:
            return r
        d = math.nextafter(d, math.inf)
    raise RuntimeError("no round-trip-exact CPP centre found")


def storm_basin(n_elem: int, seed: int, nsteps: int = 1000, p: int = 1, slopeflag: int = 5,
                wetdry: bool = True, curved: Optional[bool] = None, **kw) -> Case:
    """Spherical Delaunay basin (see delaunay_basin, ics=2) under a seeded Holland hurricane:
    central pressure U(920, 990) mbar, radius to maximum wind U(20, 60) km, translation speed
    U(4, 8) m/s toward a heading U(20, 100) degrees, on a straight track or one turning by
    U(-40, 40) degrees a day.  At the run's start the eye is inside the basin's western
    two thirds."""
    from .holland import HollandStorm, synthetic_track
    rng = np.random.default_rng(10_000 + seed)
    lon_c, lat_c = kw.pop("lon_c", -75.03), kw.pop("lat_c", 27.97)
    lx, ly = kw.get("lx", 1.0e6), kw.get("ly", 6.0e5)
    half_lon = 0.5 * lx / (R_EARTH * math.cos(math.radians(lat_c))) / DEG2RAD
    half_lat = 0.5 * ly / R_EARTH / DEG2RAD
    lon = lon_c + half_lon * rng.uniform(-0.8, 0.3)
    lat = lat_c + half_lat * rng.uniform(-0.6, 0.6)
    if curved is None:
        curved = bool(rng.integers(0, 2))
    recs = synthetic_track(rng, lon, lat, heading_deg=rng.uniform(20.0, 100.0),
                           speed_ms=rng.uniform(4.0, 8.0),
                           turn_deg_per_day=rng.uniform(-40.0, 40.0) if curved else 0.0,
                           pmin_mb=rng.uniform(920.0, 990.0), rmax_km=rng.uniform(20.0, 60.0),
                           ref=(2020, 8, 25, 0))
    storm = HollandStorm(2020, 8, 25, 0, 1, 0.9, recs)
    return delaunay_basin(n_elem=n_elem, seed=seed, p=p, nsteps=nsteps, slopeflag=slopeflag,
                          wetdry=wetdry, ics=2, lon_c=lon_c, lat_c=lat_c, storm=storm, **kw)


# ----------------------------------------------------------------------------------------------
# run directories for the reference's serial solver
# ----------------------------------------------------------------------------------------------
def _r(v: float) -> str:
    return repr(float(v))


def write_run_dir(d: str, case: Case, dghot: int = 0, dghotspool: int = 0,
                  nspoolge: Optional[int] = None) -> None:
    """Write fort.14/15/dg (+fort.13) so that dgswem_serial runs exactly `case`."""
    from .fortio import write_fort14
    os.makedirs(d, exist_ok=True)
    p = case.params
    write_fort14(os.path.join(d, "fort.14"), case, case.title)
    nt = p.nt
    rnday = nt * p.dtdp / 86400.0
    # NT = INT(RNDAY*(86400/DT)+0.5) must give back nt
    assert int(rnday * (86400.0 / p.dtdp) + 0.5) == nt
    spool = nspoolge if nspoolge is not None else (p.nspoolge or nt)
    L = []
    a = L.append
    a(case.title[:32]); a("RUN")
    a("0"); a("0"); a("1000000"); a("0")
    a(str(p.ics)); a(str(p.im)); a(str(p.nolibf)); a(str(p.nolifa)); a(str(p.nolica)); a(str(p.nolicat))
    a(str(1 if case.manning is not None else 0))
    if case.manning is not None:
        a("mannings_n_at_sea_floor")
    a(str(p.ncor)); a(str(p.ntip)); a(str(p.nws)); a(str(p.nramp))
    a(_r(p.g)); a(_r(p.tau0)); a(_r(p.dtdp)); a(_r(p.statim)); a(_r(p.reftim))
    if p.nws == 8:
        st = case.storm
        a(f"{st.irefyr} {st.irefmo} {st.irefday} {st.irefhr} {st.storm_number} {_r(st.bladj)}")
    elif p.nws != 0:
        raise ValueError("only NWS=0 and NWS=8 run directories are written")
    a(_r(rnday))
    a(_r(p.dramp)); a("0.35 0.30 0.35")
    a(_r(p.h0) if p.nolifa != 2 else f"{_r(p.h0)} 0 0 0.01")
    a(f"{_r(p.slam0 / DEG2RAD)} {_r(p.sfea0 / DEG2RAD)}")
    if p.nolibf == 0:
        a(_r(p.tau))
    elif p.nolibf == 1:
        a(_r(p.cf))
    else:
        a(f"{_r(p.cf)} {_r(p.hbreak)} {_r(p.ftheta)} {_r(p.fgamma)}")
    a(_r(p.eslm)); a(_r(p.cori))
    a(str(p.ntif))
    for i in range(p.ntif):
        a(p.tipotag[i] if i < len(p.tipotag) else f"T{i}")
        a(f"{_r(p.tpk[i])} {_r(p.amigt[i])} {_r(p.etrf[i])} {_r(p.fft[i])} {_r(p.facet[i] / DEG2RAD)}")
    a(str(p.nbfr))
    for i in range(p.nbfr):
        a(f"C{i}"); a(f"{_r(p.amig[i])} {_r(p.ff[i])} {_r(p.face[i] / DEG2RAD)}")
    for i in range(p.nbfr):
        a(f"C{i}")
        for j in range(p.emo.shape[1]):
            a(f"{_r(p.emo[i, j])} {_r(p.efa[i, j] / DEG2RAD)}")
    a("110.0")
    if any(t in (2, 12, 22) for t, _ in case.land_segs):
        a(str(p.nffr))
        for i in range(p.nffr):
            a(f"F{i}"); a(f"{_r(p.famig[i])} {_r(p.fff[i])} {_r(p.fface[i] / DEG2RAD)}")
        for i in range(p.nffr):
            a(f"F{i}")
            for j in range(p.qnam.shape[1]):
                a(f"{_r(p.qnam[i, j])} {_r(p.qnph[i, j] / DEG2RAD)}")
    a("0 0.0 1.0 0"); a("0"); a("0 0.0 1.0 0"); a("0")
    if p.nws != 0:
        a("0 0.0 1.0 0"); a("0")                      # meteorological stations
    a(f"1 0.0 {_r(rnday)} {spool}")
    a("0 0.0 1.0 0")                                  # global velocity output
    if p.nws != 0:
        a("0 0.0 1.0 0")                              # global wind output
    a("0"); a("0.0 0.0 0 0.0"); a("0 0 0 0"); a("0 0"); a("1 0 1.E-10 25")
    with open(os.path.join(d, "fort.15"), "w") as f:
        f.write("\n".join(L) + "\n")
    if p.nws == 8:
        from .holland import write_fort22
        write_fort22(os.path.join(d, "fort.22"), case.storm.records)
    if p.ntip == 2:
        with open(os.path.join(d, "fort.24"), "w") as f:
            for i in range(p.ntif):
                f.write(f"SAL {i}\n1.0\n1\n{p.tipotag[i] if i < len(p.tipotag) else 'T'}\n")
                for j in range(case.np_):
                    f.write(f"{j + 1} {_r(p.saltamp[i, j])} {_r(p.saltpha[i, j] / DEG2RAD)}\n")
    for name, ser in (("fort.19", case.elev_series), ("fort.20", case.flow_series)):
        if ser is not None:
            inc, vals = ser
            with open(os.path.join(d, name), "w") as f:
                f.write(_r(inc) + "\n")
                for v in vals.reshape(-1):
                    f.write(_r(v) + "\n")
    if case.manning is not None:
        with open(os.path.join(d, "fort.13"), "w") as f:
            f.write("attrs\n" + f"{case.np_}\n1\nmannings_n_at_sea_floor\nunitless\n1\n0.0\n")
            f.write(f"mannings_n_at_sea_floor\n{case.np_}\n")
            for i, v in enumerate(case.manning):
                f.write(f"{i + 1} {_r(v)}\n")
    g = case.dg
    with open(os.path.join(d, "fort.dg"), "w") as f:
        f.write("dgswe = 1\nmodal_ic = 0\npadapt = 0\npflag = 2\ngflag = 1\ndis_tol = 8\n")
        f.write(f"pl = {g.pl}\nph = {g.ph}\npx = {g.ph}\nslimit = 0.00005\nplimit = 10\n")
        f.write("k = 1\nks = 0.5\nL = 2\nnelem = 1\nfluxtype = 2\n")
        f.write(f"rk_stage = {g.rk_stage}\nrk_order = {g.rk_order}\n")
        f.write(f"dghot = {dghot}\ndghotspool = {dghotspool or nt}\n")
        f.write(f"slopeflag = {g.slopeflag}\nweight = 1\nsedflag = 0\nporosity = 0.0001\n")
        f.write("sevdm = 0.00001\nlayers = 1\nrxn_rate = 1\nartdif = 0\nkappa = -1.0\ns0 = 0.0\n")
        f.write(f"uniform_dif = 2.5e-6\ntune_by_hand = 0\nrainfall = {g.rainfall}\n")
