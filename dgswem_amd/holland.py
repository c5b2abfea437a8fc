"""Holland (1980) parametric hurricane forcing, NWS=8: the storm track file and its generator.

The reference reads an ATCF "best track" file (fort.22) in GetHollandStormData
(src/wind.F:2153-2450) with the fixed-column FORMAT 14 (src/wind.F:2443-2444) and evaluates the
Holland profile at every node in HollandGet (src/wind.F:1964-2137).  This module holds the file
side only: reading and writing fort.22 records, the fort.15 NWS=8 control line
(src/read_input.F:1082-1094) and a seeded synthetic track generator.  The time interpolation of the
track and the per-node wind/pressure evaluation live in the compiled library (libdgswem_hip.so)
and, for testing, in the CPU oracle.

A record is 12 integers, the FORMAT 14 fields in order:
    year, month, day, hour, cast type (0 = BEST, 1 = OFCL), forecast increment (h),
    latitude (0.1 deg N), longitude (0.1 deg W), max wind (kt), central pressure (mbar),
    radius of the last closed isobar (nm), radius to max wind (nm)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

CAST_TYPES = {"BEST": 0, "OFCL": 1}
CAST_NAMES = {v: k for k, v in CAST_TYPES.items()}

# FORMAT 14: (8x,i4,i2,i2,i2,6x,a4,2x,i3,1x,i4,3x,i4,3x,i3,2x,i4,47x,i3,2x,i3), as 0-based
# [start, end) column slices of the twelve fields
_COLS = [(8, 12), (12, 14), (14, 16), (16, 18), (24, 28), (30, 33), (34, 38), (41, 45),
         (48, 51), (53, 57), (104, 107), (109, 112)]


@dataclass
class HollandStorm:
    """NWS=8 inputs: the fort.15 line IREFYR IREFMO IREFDAY IREFHR StormNumber BLAdj and the
    fort.22 records (int32 array [n, 12])."""
    irefyr: int
    irefmo: int
    irefday: int
    irefhr: int
    storm_number: int
    bladj: float
    records: np.ndarray = field(default_factory=lambda: np.zeros((0, 12), np.int32))

    def ref(self) -> np.ndarray:
        return np.array([self.irefyr, self.irefmo, self.irefday, self.irefhr], np.int32)


def _field_int(s: str) -> int:
    # Fortran I editing with BLANK='NULL': blanks are ignored, an all-blank field reads as zero
    s = s.replace(" ", "")
    return int(s) if s else 0


def read_fort22(path: str) -> np.ndarray:
    """Records of an ATCF track file, read with the reference's FORMAT 14 columns."""
    recs = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n").ljust(112)
            if not line.strip():
                continue
            vals = []
            for k, (a, b) in enumerate(_COLS):
                s = line[a:b]
                if k == 4:
                    if s not in CAST_TYPES:
                        raise ValueError(f"{path}: invalid TECH identifier {s!r} "
                                         "(valid: BEST hindcast, OFCL forecast)")
                    vals.append(CAST_TYPES[s])
                else:
                    vals.append(_field_int(s))
            recs.append(vals)
    return np.array(recs, dtype=np.int32).reshape(-1, 12)


def format_record(r) -> str:
    """One fort.22 line whose FORMAT 14 fields read back as `r`."""
    yr, mo, dy, hr, ct, inc, lat, lon, spd, cp, rrp, rmw = (int(v) for v in r)
    head = f"AL, 01, {yr:04d}{mo:02d}{dy:02d}{hr:02d},   , {CAST_NAMES[ct]}, {inc:3d},{lat:4d}N, {lon:4d}W, "
    head += f"{spd:3d}, {cp:4d}"
    mid = ", XX,  34, NEQ,    0,    0,    0,    0, 1013, "
    line = head + mid.ljust(47) + f"{rrp:3d}, {rmw:3d}, "
    assert len(head) == 57 and line[104:107] == f"{rrp:3d}"
    return line


def write_fort22(path: str, records: np.ndarray) -> None:
    with open(path, "w") as f:
        for r in np.asarray(records):
            f.write(format_record(r) + "\n")


def timeconv(iyr: int, imo: int, iday: int, ihr: int) -> float:
    """Seconds since the start of the year (src/wind.F:315-348, minutes and seconds zero)."""
    t = float((iday - 1) * 86400 + ihr * 3600)
    days = [31, 29 if (iyr // 4) * 4 == iyr else 28, 31, 30, 31, 30, 31, 31, 30, 31, 30]
    for m in range(2, 13):
        if imo >= m:
            t += days[m - 2] * 86400
    return t


def _add_hours(yr, mo, dy, hr, h):
    import datetime
    t = datetime.datetime(yr, mo, dy, hr) + datetime.timedelta(hours=h)
    return t.year, t.month, t.day, t.hour


def _move(lon, lat, heading_deg, dist_m):
    lat2 = lat + dist_m * math.cos(math.radians(heading_deg)) / 111195.0
    lon2 = lon + dist_m * math.sin(math.radians(heading_deg)) / (111195.0 * math.cos(math.radians(lat)))
    return lon2, lat2


def synthetic_track(rng: np.random.Generator, lon: float, lat: float, heading_deg: float,
                    speed_ms: float, turn_deg_per_day: float, pmin_mb: float, rmax_km: float,
                    ref=(2020, 8, 25, 0), hours_before: int = 24, hours_after: int = 72,
                    every: int = 6) -> np.ndarray:
    """BEST records every `every` hours of a storm that is at (lon, lat) (degrees, east positive)
    at the reference date `ref`, moving at `speed_ms` toward `heading_deg` (0 = north, 90 = east)
    and turning by `turn_deg_per_day` (0 for a straight track).  Maximum wind from the central
    pressure by the Atkinson-Holliday relation V = 6.7 (1010 - P)^0.644 kt; the radius of the
    last closed isobar is 6 Rmax.  Positions are rounded to 0.1 degree and intensities to
    integers, as the file format holds them."""
    step = speed_ms * every * 3600.0
    turn = turn_deg_per_day * every / 24.0
    pts = {0: (lon, lat, heading_deg)}
    lo, la, hd = lon, lat, heading_deg
    for k in range(1, hours_after // every + 1):
        lo, la = _move(lo, la, hd, step)
        hd += turn
        pts[k] = (lo, la, hd)
    lo, la, hd = lon, lat, heading_deg
    for k in range(1, hours_before // every + 1):
        hd -= turn
        lo, la = _move(lo, la, hd + 180.0, step)
        pts[-k] = (lo, la, hd)
    rmw = max(1, int(round(rmax_km / 1.852)))
    recs = []
    for k in sorted(pts):
        lo, la, _ = pts[k]
        yr, mo, dy, hr = _add_hours(*ref, k * every)
        p = int(round(pmin_mb + rng.uniform(-2.0, 2.0)))
        vmax = int(round(6.7 * max(1010.0 - p, 1.0) ** 0.644))
        recs.append([yr, mo, dy, hr, 0, 0, int(round(la * 10.0)), int(round(-lo * 10.0)), vmax, p,
                     6 * rmw, rmw])
    return np.array(recs, dtype=np.int32)


def forecast_track(best: np.ndarray, issue: int, incs=(0, 12, 12, 24, 36, 48)) -> np.ndarray:
    """A BEST/OFCL mixture built from a 6-hourly BEST track: BEST lines up to index `issue`, each
    written twice (as ATCF repeats a time for each wind-radius quadrant), then an advisory issued at
    that time whose OFCL lines carry forecast increments `incs` (hours; repeats included) with the
    positions and winds of the BEST records that many hours later and a central pressure and radii
    that the reader must replace by the nowcast's."""
    recs = []
    for r in best[:issue + 1]:
        recs.append(list(r)); recs.append(list(r))
    base = best[issue]
    for inc in incs:
        src = best[min(issue + inc // 6, len(best) - 1)]
        recs.append([base[0], base[1], base[2], base[3], 1, inc, src[6], src[7], src[8],
                     1005, 50, 40])
    return np.array(recs, dtype=np.int32)
