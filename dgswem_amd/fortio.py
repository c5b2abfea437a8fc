"""ADCIRC/DG-SWEM input and output files: fort.14, fort.15, fort.13, fort.dg, fort.63, DG hot start.

This is the host-side mirror of the reference's input stage for the explicit hydro step: it reads
the same files `dgswem_serial` reads and derives the same per-node quantities `READ_INPUT` derives
(projected coordinates, spherical factor, depth floor, Coriolis, tidal periods, number of steps), so
that the GPU step sees bit-identical inputs.  Only the features that feed the hot path are parsed;
everything after the global-output lines of fort.15 (harmonic analysis, hot-start spooling, solver
parameters) is ignored, as are station coordinates.

Reference files followed:
  src/read_input.F:209-3931   fort.15 record order (list-directed: each READ starts a new record)
  src/read_input.F:1284-1420  fort.14 nodes/elements, CPP projection, SFAC, DP floor, AREAS sign check
  src/read_input.F:1660-1860  friction line, ESLM, CORI/CORIF
  src/read_input.F:1979-2111  NBFR constituents, open boundaries, EMO/EFA
  src/read_input.F:2240-2440  land boundary segments (types 0,1,2,10-13,20-25,30)
  src/read_input.F:2960-3045  flux boundary constituents QNAM/QNPH (DGSWE=1 ordering)
  src/read_input.F:3165-3935  station and global output specs
  src/read_fort_dg.F90:73-185,449-484  fort.dg fixed and keyword formats
  src/nodalattr.F:295-739     fort.13 nodal attributes (Manning's n)
  src/dgswem.F:614-640        CPP / INVCP projections
  src/write_results.F:299-372 fort.63 ASCII layout (E16.8E3, -99999 for dry nodes)
  src/write_results.F:1199-1259  DG hot start Hot_start.263/264/214
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

PI = 3.141592653589793          # src/global.F:72
DEG2RAD = PI / 180.0            # src/global.F:73
R_EARTH = 6378206.4             # src/dgswem.F:617

# Land-boundary segment types that become no-normal-flow edges and flux-specified edges
# (src/create_edge_data.F:391-405 and the classification that follows).
LAND_TYPES = (0, 10, 20, 1, 11, 21)
FLOW_TYPES = (2, 12, 22)


@dataclass
class Params:
    """Run-control scalars after READ_INPUT's derivations (names follow the reference)."""
    ics: int = 1
    im: int = 0
    nolibf: int = 1
    nolifa: int = 1
    nolica: int = 1
    nolicat: int = 1
    nwp: int = 0
    ncor: int = 0
    ntip: int = 0
    nws: int = 0
    nramp: int = 1
    g: float = 9.81
    tau0: float = 0.005
    dtdp: float = 1.0
    statim: float = 0.0
    reftim: float = 0.0
    rnday: float = 1.0
    dramp: float = 1.0
    h0: float = 0.01
    slam0: float = 0.0            # radians
    sfea0: float = 0.0            # radians
    tau: float = 0.0
    cf: float = 0.0
    hbreak: float = 1.0
    ftheta: float = 1.0
    fgamma: float = 1.0
    eslm: float = 0.0
    cori: float = 0.0
    nt: int = 0
    # open-boundary constituents
    nbfr: int = 0
    amig: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ff: np.ndarray = field(default_factory=lambda: np.zeros(0))
    face: np.ndarray = field(default_factory=lambda: np.zeros(0))   # radians
    per: np.ndarray = field(default_factory=lambda: np.zeros(0))
    emo: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))  # [nbfr, neta]
    efa: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))  # radians
    # flux-boundary constituents
    nffr: int = 0
    famig: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fff: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fface: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fper: np.ndarray = field(default_factory=lambda: np.zeros(0))
    qnam: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))  # [nffr, n flux nodes]
    qnph: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    # tidal potential constituents (NTIP = 1, 2; src/read_input.F:1856-1923)
    ntif: int = 0
    tpk: np.ndarray = field(default_factory=lambda: np.zeros(0))
    amigt: np.ndarray = field(default_factory=lambda: np.zeros(0))
    etrf: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fft: np.ndarray = field(default_factory=lambda: np.zeros(0))
    facet: np.ndarray = field(default_factory=lambda: np.zeros(0))   # radians
    pert: np.ndarray = field(default_factory=lambda: np.zeros(0))
    tipotag: List[str] = field(default_factory=list)
    saltamp: Optional[np.ndarray] = None     # [ntif, np] (NTIP = 2, fort.24)
    saltpha: Optional[np.ndarray] = None     # [ntif, np] radians
    # global elevation output (fort.63)
    noutge: int = 0
    toutsge: float = 0.0
    toutfge: float = 0.0
    nspoolge: int = 0


@dataclass
class DGOptions:
    """fort.dg keys used by the hot path (src/read_fort_dg.F90:449-484 defaults)."""
    pl: int = 1
    ph: int = 1
    fluxtype: int = 1
    rk_stage: int = 2
    rk_order: int = 2
    modal_ic: int = 0
    dghot: int = 0
    dghotspool: int = 86400
    slopeflag: int = 5
    artdif: int = 0
    rainfall: int = 0
    padapt: int = 0


@dataclass
class Case:
    """Everything the explicit hydro step needs, in the reference's conventions (1-based NM)."""
    title: str
    params: Params
    dg: DGOptions
    x: np.ndarray            # projected x of nodes (m)
    y: np.ndarray            # projected y of nodes (m)
    slam: np.ndarray         # longitude (rad)
    sfea: np.ndarray         # latitude (rad)
    dp: np.ndarray           # depth after the H0 floor (NOLIFA<2)
    sfac: np.ndarray         # spherical factor per node
    corif: np.ndarray        # Coriolis parameter per node
    nm: np.ndarray           # [NE,3] int32, 1-based, counter-clockwise
    open_segs: List[np.ndarray]                  # NBDV per open boundary segment (1-based)
    land_segs: List[Tuple[int, np.ndarray]]      # (IBTYPE, NBVV) per land segment (1-based)
    barrier_back: dict = field(default_factory=dict)  # segment index -> IBCONN (types 4/24/5/25)
    manning: Optional[np.ndarray] = None         # ManningsN per node when loaded from fort.13
    storm: Optional["HollandStorm"] = None        # NWS=8 track and control line
    lonlat: Optional[np.ndarray] = None          # [NP,2] fort.14 lon/lat in degrees (ICS=2)
    # non-periodic boundary forcing: fort.19 (NBFR = 0 with open boundaries) as
    # (ETIMINC, [nrec, NETA]) and fort.20 (NFFR = 0 with flux boundaries) as (FTIMINC, [nrec, flux nodes])
    elev_series: Optional[Tuple[float, np.ndarray]] = None
    flow_series: Optional[Tuple[float, np.ndarray]] = None

    def lbcodei(self) -> np.ndarray:
        """LBCODEI(NVEL): the boundary type of every land-boundary node, in NVEL order."""
        return np.concatenate([np.full(len(n), t, np.int32) for t, n in self.land_segs]) \
            if self.land_segs else np.zeros(0, np.int32)

    def nbd(self) -> np.ndarray:
        """NBD(NETA): the open-boundary nodes in order (1-based)."""
        return np.concatenate(self.open_segs).astype(np.int32) if self.open_segs \
            else np.zeros(0, np.int32)

    @property
    def ne(self) -> int:
        return int(self.nm.shape[0])

    @property
    def np_(self) -> int:
        return int(self.x.shape[0])

    @property
    def dofh(self) -> int:
        p = self.dg.ph
        return (p + 1) * (p + 2) // 2


# ----------------------------------------------------------------------------------------------
# list-directed record reader
# ----------------------------------------------------------------------------------------------
_SPLIT = re.compile(r"[,\s]+")


class _Records:
    """Sequential reader reproducing Fortran list-directed READ(u,*) record semantics."""

    def __init__(self, path: str):
        with open(path, "r", newline="") as f:
            text = f.read()
        # the reference fixtures use '\n' as the record separator (some headers carry stray '\r')
        self.lines = text.split("\n")
        self.i = 0

    def text(self) -> str:
        s = self.lines[self.i]
        self.i += 1
        return s.rstrip("\r")

    def values(self, n: int) -> List[str]:
        out: List[str] = []
        while len(out) < n:
            if self.i >= len(self.lines):
                raise EOFError("unexpected end of file")
            line = self.lines[self.i]
            self.i += 1
            # a '!' starts a trailing comment in the reference's input files; list-directed
            # input stops consuming once it has all values, so the comment is never parsed.
            line = line.split("!")[0]
            toks = [t for t in _SPLIT.split(line.strip()) if t]
            if "/" in toks:
                toks = toks[: toks.index("/")]
            out.extend(toks[: n - len(out)])
        return out

    def ints(self, n: int = 1):
        v = [int(float(t)) if ("." in t or "e" in t.lower()) else int(t) for t in self.values(n)]
        return v if n > 1 else v[0]

    def reals(self, n: int = 1):
        v = [_freal(t) for t in self.values(n)]
        return v if n > 1 else v[0]


def _freal(t: str) -> float:
    return float(t.replace("D", "E").replace("d", "e"))


# ----------------------------------------------------------------------------------------------
# fort.dg
# ----------------------------------------------------------------------------------------------
def read_fort_dg(path: str) -> DGOptions:
    """Fixed format if the first token is '1' (src/read_input.F:193-201), else keyword format."""
    o = DGOptions()
    with open(path) as f:
        lines = f.read().split("\n")
    first = next((l for l in lines if l.strip()), "").strip()
    if first.startswith("1"):
        r = _Records(path)
        r.ints(1)                                  # DGSWE
        o.padapt = r.ints(2)[0]                    # padapt, pflag
        r.values(2)                                # gflag, diorism
        o.pl, o.ph, _ = r.ints(3)
        r.values(1); r.values(1); r.values(3)      # slimit, plimit, pflag2con1/2, lebesgueP
        o.fluxtype = r.ints(1)
        o.rk_stage, o.rk_order = r.ints(2)
        r.values(1)                                # DG_TO_CG
        o.modal_ic = r.ints(1)
        o.dghot, o.dghotspool = r.ints(2)
        o.slopeflag = int(r.values(1)[0])
        r.values(4); r.values(1); r.values(1)      # sedflag..layers, reaction_rate, MNES
        o.artdif = r.ints(5)[0]
        return o
    keys = {}
    for line in lines:
        s = line.split("!")[0].strip()
        if not s or "=" not in s:
            continue
        k, v = s.split("=", 1)
        keys[k.strip()] = v.strip()
    ival = lambda k, d: int(float(_freal(keys[k]))) if k in keys else d
    o.pl = ival("pl", o.pl)
    o.ph = ival("ph", o.ph)
    o.padapt = ival("padapt", o.padapt)
    o.fluxtype = ival("fluxtype", o.fluxtype)
    o.rk_stage = ival("rk_stage", o.rk_stage)
    o.rk_order = ival("rk_order", o.rk_order)
    o.modal_ic = ival("modal_ic", o.modal_ic)
    o.dghot = ival("dghot", o.dghot)
    o.dghotspool = ival("dghotspool", o.dghotspool)
    o.slopeflag = ival("slopeflag", o.slopeflag)
    o.artdif = ival("artdif", o.artdif)
    o.rainfall = ival("rainfall", o.rainfall)
    return o


# ----------------------------------------------------------------------------------------------
# fort.13 (only the Manning's n attribute is used by the hot path)
# ----------------------------------------------------------------------------------------------
def read_fort13(path: str, names: List[str], n_nodes: int) -> dict:
    r = _Records(path)
    r.text()                                        # header
    nn = r.ints(1)
    if nn != n_nodes:
        raise ValueError(f"fort.13 has {nn} nodes, fort.14 has {n_nodes}")
    nattr = r.ints(1)
    defaults = {}
    for _ in range(nattr):
        name = r.text().strip()
        r.text()                                    # units
        nvals = r.ints(1)
        dv = r.reals(nvals)
        defaults[name] = (nvals, dv if nvals > 1 else [dv])
    out = {}
    for _ in range(nattr):
        name = r.text().strip()
        nnd = r.ints(1)
        nvals, dv = defaults[name]
        if nvals == 1:
            arr = np.full(n_nodes, dv[0], dtype=np.float64)
            for _k in range(nnd):
                t = r.values(2)
                arr[int(t[0]) - 1] = _freal(t[1])
        else:
            arr = np.tile(np.asarray(dv, dtype=np.float64), (n_nodes, 1))
            for _k in range(nnd):
                t = r.values(1 + nvals)
                arr[int(t[0]) - 1, :] = [_freal(s) for s in t[1:]]
        if name in names:
            out[name] = arr
    return out


# ----------------------------------------------------------------------------------------------
# fort.14 + fort.15
# ----------------------------------------------------------------------------------------------
def read_case(dirname: str) -> Case:
    """Read fort.14/fort.15/fort.dg (+fort.13 when NWP>0) from a run directory."""
    dg = read_fort_dg(os.path.join(dirname, "fort.dg"))
    # with padapt=0 the reference runs every element at order pl (src/prep_DG.F:90,104,124); the
    # step here runs one order, ph, so the two must agree
    if dg.padapt != 0:
        raise NotImplementedError(f"fort.dg padapt={dg.padapt}: p-adaptivity is not supported")
    if dg.pl != dg.ph:
        raise NotImplementedError(f"fort.dg pl={dg.pl} differs from ph={dg.ph}: mixed orders are "
                                  "not supported")
    if dg.artdif != 0:
        raise NotImplementedError(f"fort.dg artdif={dg.artdif}: LDG artificial diffusion is not "
                                  "supported")
    f15 = _Records(os.path.join(dirname, "fort.15"))
    f14 = _Records(os.path.join(dirname, "fort.14"))
    p = Params()
    title = f15.text()[:32].strip()
    f15.text()                                        # RUNID
    f15.ints(1); f15.ints(1); f15.ints(1)             # NFOVER, NABOUT, NSCREEN
    ihot = f15.ints(1)
    if ihot != 0:
        raise NotImplementedError("IHOT != 0 (ADCIRC hot start) is not part of the DG step path")
    p.ics = f15.ints(1)
    p.im = f15.ints(1)
    p.nolibf = f15.ints(1)
    p.nolifa = f15.ints(1)
    p.nolica = f15.ints(1)
    p.nolicat = f15.ints(1)
    p.nwp = f15.ints(1)
    attr_names = [f15.text().strip() for _ in range(p.nwp)] if p.nwp > 0 else []
    p.ncor = f15.ints(1)
    p.ntip = f15.ints(1)
    p.nws = f15.ints(1)
    if p.nws not in (0, 8):
        raise NotImplementedError("only NWS=0 and NWS=8 (Holland vortex) meteorological forcing "
                                  "are supported")
    p.nramp = f15.ints(1)
    if p.nramp > 1:
        raise NotImplementedError("NRAMP > 1")
    p.g = f15.reals(1)
    p.tau0 = f15.reals(1)
    p.dtdp = f15.reals(1)
    p.statim = f15.reals(1)
    p.reftim = f15.reals(1)
    storm = None
    if p.nws == 8:                                     # src/read_input.F:1082-1094
        t = f15.values(6)
        from .holland import HollandStorm, read_fort22
        storm = HollandStorm(int(t[0]), int(t[1]), int(t[2]), int(t[3]), int(t[4]), _freal(t[5]))
        storm.records = read_fort22(os.path.join(dirname, "fort.22"))
    p.rnday = f15.reals(1)
    p.nt = int(p.rnday * (86400.0 / p.dtdp) + 0.5)      # src/read_input.F:1156
    p.dramp = f15.reals(1)
    f15.reals(3)                                       # A00 B00 C00
    if p.nolifa != 2:
        p.h0 = f15.reals(1)
    else:
        p.h0 = f15.reals(4)[0]

    # ---- fort.14 nodes & elements (src/read_input.F:1284-1420)
    f14.text()
    ne, n_nodes = f14.ints(2)
    slam0, sfea0 = f15.reals(2)
    p.slam0, p.sfea0 = slam0 * DEG2RAD, sfea0 * DEG2RAD
    x = np.empty(n_nodes); y = np.empty(n_nodes); dp = np.empty(n_nodes)
    slam = np.empty(n_nodes); sfea = np.empty(n_nodes)
    lonlat = np.empty((n_nodes, 2)) if p.ics == 2 else None
    for i in range(n_nodes):
        t = f14.values(4)
        j = int(t[0]) - 1
        a, b, d = _freal(t[1]), _freal(t[2]), _freal(t[3])
        dp[j] = d
        if p.ics == 1:
            x[j], y[j] = a, b
            slam[j] = p.slam0 + a / (R_EARTH * math.cos(p.sfea0))       # INVCP
            sfea[j] = b / R_EARTH
        else:
            lonlat[j] = a, b
    if p.ics == 1:
        sfac = np.ones(n_nodes)
    else:
        slam, sfea, x, y, sfac = spherical_coords(lonlat, p.slam0, p.sfea0)
    if p.nolifa in (0, 1):
        dp = np.where(dp < p.h0, p.h0, dp)
    nm = np.empty((ne, 3), dtype=np.int32)
    for i in range(ne):
        t = f14.values(5)
        j = int(t[0]) - 1
        nm[j] = [int(t[2]), int(t[3]), int(t[4])]
    x1, x2, x3 = x[nm[:, 0] - 1], x[nm[:, 1] - 1], x[nm[:, 2] - 1]
    y1, y2, y3 = y[nm[:, 0] - 1], y[nm[:, 1] - 1], y[nm[:, 2] - 1]
    areas = (x1 - x3) * (y2 - y3) + (x3 - x2) * (y1 - y3)
    if np.any(areas < 0.0):
        bad = int(np.nonzero(areas < 0.0)[0][0]) + 1
        raise ValueError(f"element {bad} is not counter-clockwise")

    # ---- friction / viscosity / Coriolis (src/read_input.F:1660-1840)
    if p.nolibf == 0:
        p.tau = f15.reals(1); p.cf = p.tau
    elif p.nolibf == 1:
        p.cf = f15.reals(1)
    else:
        p.cf, p.hbreak, p.ftheta, p.fgamma = f15.reals(4)
    p.eslm = f15.reals(2)[0] if p.im == 10 else f15.reals(1)
    manning = None
    if p.nwp > 0:
        if "quadratic_friction_coefficient_at_sea_floor" in attr_names:
            raise NotImplementedError("a per-node quadratic friction attribute (FRIC): the step "
                                      "takes CF uniform")
        attrs = read_fort13(os.path.join(dirname, "fort.13"), attr_names, n_nodes)
        manning = attrs.get("mannings_n_at_sea_floor")
    p.cori = f15.reals(1)
    if p.ncor == 0:
        corif = np.full(n_nodes, p.cori)
    else:
        corif = coriolis(sfea)

    # ---- tidal potential constituents (src/read_input.F:1856-1923)
    p.ntif = f15.ints(1)
    cons = []
    for _ in range(p.ntif):
        p.tipotag.append(f15.text()[:5].strip())
        cons.append(f15.reals(5))
    cons = np.array(cons, dtype=np.float64).reshape(p.ntif, 5)
    p.tpk, p.amigt, p.etrf, p.fft = cons[:, 0].copy(), cons[:, 1].copy(), cons[:, 2].copy(), cons[:, 3].copy()
    p.facet = cons[:, 4] * DEG2RAD
    p.pert = np.array([0.0 if a == 0.0 else 2.0 * PI / a for a in p.amigt])
    if p.ntip not in (0, 1, 2):
        raise NotImplementedError(f"NTIP={p.ntip}")
    if p.ntip == 2:
        p.saltamp, p.saltpha = read_fort24(os.path.join(dirname, "fort.24"), p.ntif, n_nodes)

    # ---- open-boundary constituents and segments (src/read_input.F:1979-2111)
    p.nbfr = f15.ints(1)
    amig, ff, face = [], [], []
    for _ in range(p.nbfr):
        f15.text()
        a, b, c = f15.reals(3)
        amig.append(a); ff.append(b); face.append(c * DEG2RAD)
    p.amig = np.array(amig); p.ff = np.array(ff); p.face = np.array(face)
    p.per = np.array([0.0 if a == 0.0 else 2.0 * PI / a for a in amig])
    nope = f14.ints(1)
    neta = f14.ints(1)
    open_segs = []
    for _ in range(nope):
        nv = f14.ints(1)
        open_segs.append(np.array([f14.ints(1) for _ in range(nv)], dtype=np.int32))
    neta_real = sum(len(s) for s in open_segs)
    emo = np.zeros((p.nbfr, neta_real)); efa = np.zeros((p.nbfr, neta_real))
    for i in range(p.nbfr):
        f15.text()
        for j in range(neta_real):
            a, b = f15.reals(2)
            emo[i, j] = a; efa[i, j] = b * DEG2RAD
    p.emo, p.efa = emo, efa
    f15.reals(1)                                      # ANGINN

    # ---- land boundary segments (src/read_input.F:2240-2440)
    nbou = f14.ints(1)
    f14.ints(1)                                       # NVEL
    land_segs = []
    barrier_back = {}
    for k in range(nbou):
        nvell, ibtype = f14.ints(2)
        nodes, back = [], []
        for _k in range(nvell):
            if ibtype in (3, 13, 23):
                nodes.append(int(f14.values(3)[0]))
            elif ibtype in (4, 24, 5, 25):
                t = f14.values(5 if ibtype in (4, 24) else 4)
                nodes.append(int(t[0])); back.append(int(t[1]))
            else:
                nodes.append(f14.ints(1))
        land_segs.append((ibtype, np.array(nodes, dtype=np.int32)))
        if back:
            barrier_back[k] = np.array(back, dtype=np.int32)

    # ---- flux boundary constituents (src/read_input.F:2960-3045, DGSWE=1 ordering)
    if any(t in FLOW_TYPES for t, _ in land_segs):
        p.nffr = f15.ints(1)
        famig, fff, fface = [], [], []
        for _ in range(p.nffr):
            f15.text()
            a, b, c = f15.reals(3)
            famig.append(a); fff.append(b); fface.append(c * DEG2RAD)
        p.famig = np.array(famig); p.fff = np.array(fff); p.fface = np.array(fface)
        p.fper = np.array([0.0 if a == 0.0 else 2.0 * PI / a for a in famig])
        nflow_nodes = sum(len(n) for t, n in land_segs if t in FLOW_TYPES)
        qnam = np.zeros((max(p.nffr, 1), nflow_nodes)); qnph = np.zeros_like(qnam)
        for i in range(p.nffr):
            f15.text()
            for j in range(nflow_nodes):
                a, b = f15.reals(2)
                qnam[i, j] = a; qnph[i, j] = b * DEG2RAD
        p.qnam, p.qnph = qnam, qnph

    # ---- non-periodic boundary forcing files (src/cstart.F:80-124)
    elev_series = flow_series = None
    if open_segs and p.nbfr == 0:
        elev_series = read_series(os.path.join(dirname, "fort.19"), sum(len(s) for s in open_segs))
    nflux = sum(len(n) for t, n in land_segs if t in FLOW_TYPES)
    if nflux and p.nffr == 0:
        flow_series = read_series(os.path.join(dirname, "fort.20"), nflux)

    # ---- station output specs (skipped) and global elevation output
    for _ in range(2):                                # elevation, velocity stations
        f15.values(4)
        nsta = f15.ints(1)
        for _k in range(nsta):
            f15.reals(2)
    if p.im == 10:
        f15.values(4)
        nsta = f15.ints(1)
        for _k in range(nsta):
            f15.reals(2)
    if p.nws != 0:
        f15.values(4)
        nsta = f15.ints(1)
        for _k in range(nsta):
            f15.reals(2)
    t = f15.values(4)
    p.noutge, p.toutsge, p.toutfge, p.nspoolge = int(t[0]), _freal(t[1]), _freal(t[2]), int(t[3])

    return Case(title=title, params=p, dg=dg, x=x, y=y, slam=slam, sfea=sfea, dp=dp, sfac=sfac,
                corif=corif, nm=nm, open_segs=open_segs, land_segs=land_segs, manning=manning,
                barrier_back=barrier_back, storm=storm, lonlat=lonlat, elev_series=elev_series,
                flow_series=flow_series)


def read_series(path: str, nvals: int) -> Tuple[float, np.ndarray]:
    """fort.19 / fort.20: the time increment, then records of `nvals` values, one per line
    (list-directed READs, src/cstart.F:89-94,109-121, src/DG_hydro_timestep.F:158,181)."""
    r = _Records(path)
    inc = r.reals(1)
    vals = []
    while True:
        try:
            vals.append(r.reals(1))
        except (EOFError, IndexError, ValueError):
            break
    nrec = len(vals) // nvals
    if nrec < 2:
        raise ValueError(f"{path}: fewer than two records of {nvals} values")
    return inc, np.array(vals[:nrec * nvals], dtype=np.float64).reshape(nrec, nvals)


def read_fort24(path: str, ntif: int, n_nodes: int) -> Tuple[np.ndarray, np.ndarray]:
    """fort.24 self-attraction/load tide (src/read_input.F:1880-1889): per constituent a
    four-line header, then JJ, SALTAMP, SALTPHA (degrees) for every node."""
    r = _Records(path)
    amp = np.zeros((ntif, n_nodes)); pha = np.zeros((ntif, n_nodes))
    for i in range(ntif):
        for _ in range(4):
            r.text()
        for _ in range(n_nodes):
            t = r.values(3)
            jj = int(t[0])
            amp[i, jj - 1] = _freal(t[1])
            pha[i, jj - 1] = _freal(t[2]) * DEG2RAD
    return amp, pha


def spherical_coords(lonlat: np.ndarray, slam0: float, sfea0: float):
    """ICS=2 node coordinates from fort.14 lon/lat degrees: radians, the CPP projection and
    SFAC (src/read_input.F:1322-1348)."""
    slam = np.empty(len(lonlat)); sfea = np.empty(len(lonlat))
    x = np.empty(len(lonlat)); y = np.empty(len(lonlat)); sfac = np.empty(len(lonlat))
    c0 = math.cos(sfea0)
    for j, (a, b) in enumerate(lonlat):
        slam[j] = DEG2RAD * float(a)
        sfea[j] = DEG2RAD * float(b)
        x[j] = R_EARTH * (slam[j] - slam0) * c0
        y[j] = sfea[j] * R_EARTH
        sfac[j] = c0 / math.cos(sfea[j])
    return slam, sfea, x, y, sfac


def coriolis(sfea: np.ndarray) -> np.ndarray:
    """NCOR=1 Coriolis parameter per node (src/read_input.F:1825-1830)."""
    return np.array([2.0 * 7.29212e-5 * math.sin(s) for s in sfea])


BARRIER_TYPES = (3, 13, 23, 4, 24, 5, 25)


def barriers_as_land(case: Case) -> Case:
    """Return a copy whose barrier segments are plain no-normal-flow segments.

    Barrier edges (src/ebarrier_edge_hydro.F, src/ibarrier_edge_hydro.F) lie outside the hot path.
    Below the barrier height they reflect the normal flow exactly like a land edge; an internal
    barrier's back node string (IBCONN) becomes a land segment of its own.
    """
    import copy
    c = copy.copy(case)
    segs = []
    for k, (t, nodes) in enumerate(case.land_segs):
        if t in BARRIER_TYPES:
            segs.append((0, nodes.copy()))
            if k in case.barrier_back:
                segs.append((0, case.barrier_back[k].copy()))
        else:
            segs.append((t, nodes))
    c.land_segs = segs
    c.barrier_back = {}
    return c


def write_fort14(path: str, case: Case, title: str = "grid") -> None:
    """Write a fort.14 (ICS=1 coordinates or lon/lat in degrees for ICS=2)."""
    with open(path, "w") as f:
        f.write(title[:24] + "\n")
        f.write(f"{case.ne} {case.np_}\n")
        for i in range(case.np_):
            if case.params.ics == 1:
                a, b = case.x[i], case.y[i]
            elif case.lonlat is not None:
                a, b = case.lonlat[i]
            else:
                a, b = case.slam[i] / DEG2RAD, case.sfea[i] / DEG2RAD
            f.write(f"{i + 1} {float(a)!r} {float(b)!r} {float(case.dp[i])!r}\n")
        for e in range(case.ne):
            n = case.nm[e]
            f.write(f"{e + 1} 3 {n[0]} {n[1]} {n[2]}\n")
        f.write(f"{len(case.open_segs)}\n{sum(len(s) for s in case.open_segs)}\n")
        for s in case.open_segs:
            f.write(f"{len(s)}\n")
            for n in s:
                f.write(f"{n}\n")
        f.write(f"{len(case.land_segs)}\n{sum(len(s) for _, s in case.land_segs)}\n")
        for t, s in case.land_segs:
            if t in BARRIER_TYPES:
                raise ValueError("write_fort14 writes plain boundary types only")
            f.write(f"{len(s)} {t}\n")
            for n in s:
                f.write(f"{n}\n")


# ----------------------------------------------------------------------------------------------
# output files
# ----------------------------------------------------------------------------------------------
def output_steps(p: Params) -> List[int]:
    """Steps at which fort.63 is written (src/read_input.F:3865-3900, write_results.F:299-302)."""
    if p.noutge == 0 or p.nspoolge == 0:
        return []
    ntcys = int((p.toutsge - p.statim) * (86400.0 / p.dtdp) + 0.5)
    ntcyf = min(int((p.toutfge - p.statim) * (86400.0 / p.dtdp) + 0.5), p.nt)
    steps, cnt = [], 0
    for it in range(1, p.nt + 1):
        if ntcys < it <= ntcyf:
            cnt += 1
            if cnt == p.nspoolge:
                steps.append(it)
                cnt = 0
    return steps


def _fmt_e16_8e3(v: float) -> str:
    """Fortran E16.8E3: ' -0.21064503E+000' style."""
    if v == 0.0:
        return "  0.00000000E+000"
    s = "-" if v < 0 else ""
    a = abs(v)
    e = int(math.floor(math.log10(a))) + 1
    m = a / 10.0 ** e
    mant = round(m, 8)
    if mant >= 1.0:
        e += 1
        mant = round(a / 10.0 ** e, 8)
    if mant < 0.1:
        e -= 1
        mant = round(a / 10.0 ** e, 8)
    es = "+" if e >= 0 else "-"
    body = f"{s}0.{int(round(mant * 1e8)):08d}E{es}{abs(e):03d}"
    return body.rjust(16)


def write_fort63_record(f, time_a: float, it: int, eta: np.ndarray,
                        nodecode: Optional[np.ndarray], nolifa: int) -> None:
    """One fort.63 snapshot (src/write_results.F:303-331)."""
    f.write(f"  {time_a:20.10E}     {it:10d}\n")
    for i, v in enumerate(eta):
        if abs(v) <= 1e-30:
            v = 0.0
        if nolifa == 2 and nodecode is not None and nodecode[i] == 0:
            v = -99999.0
        f.write(f"  {i + 1:8d}  {_fmt_e16_8e3(float(v))}\n")


def read_fort63(path: str):
    """Parse an ASCII fort.63 into (times, steps, [nsnap, np] values)."""
    with open(path, newline="") as f:
        lines = f.read().split("\n")
    hdr = lines[1].split()
    nrec, n_nodes = int(hdr[0]), int(hdr[1])
    times, steps, data = [], [], []
    i = 2
    while i < len(lines) and lines[i].strip():
        t = lines[i].split()
        times.append(_freal(t[0])); steps.append(int(t[1]))
        i += 1
        data.append([_freal(l.split()[1]) for l in lines[i:i + n_nodes]])
        i += n_nodes
    return np.array(times), np.array(steps), np.array(data)


def read_dg_hotstart(dirname: str, dofh: int, ne: int):
    """Read Hot_start.263/264/214 (list-directed, element-major, modes inner)."""
    def nums(path):
        with open(path) as f:
            return [t for t in _SPLIT.split(f.read().strip()) if t]
    z = nums(os.path.join(dirname, "Hot_start.263"))
    q = nums(os.path.join(dirname, "Hot_start.264"))
    h = nums(os.path.join(dirname, "Hot_start.214"))
    ze = np.array([_freal(t) for t in z[1:1 + dofh * ne]]).reshape(ne, dofh)
    qq = np.array([_freal(t) for t in q[2:2 + 2 * dofh * ne]]).reshape(ne, dofh, 2)
    it = int(h[0])
    hw = h[2:2 + 2 * dofh * ne]
    wd = np.array([int(float(hw[2 * k + 1])) for k in range(0, dofh * ne, dofh)], dtype=np.int32)
    hb = np.array([_freal(hw[2 * k]) for k in range(dofh * ne)]).reshape(ne, dofh)
    return {"it": it, "ze": ze, "qx": qq[:, :, 0].copy(), "qy": qq[:, :, 1].copy(),
            "hb": hb, "wdflg": wd}
