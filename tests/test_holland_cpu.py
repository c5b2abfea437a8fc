"""NWS=8 Holland vortex forcing on the CPU: the fort.22 reader, the track processing of the
compiled library (host code, no GPU) against the oracle's, and the oracle's node forcing against
closed forms of the Holland (1980) profile.

The oracle's NWS=8 path is pinned bit for bit against dgswem_serial by the `storm3k` golden states
(test_oracle_golden.py); these tests pin the pieces separately so that a wrong term shows up here
with its name.
"""
import datetime
import math

import numpy as np
import pytest

from conftest import load_case
from dgswem_amd import gpu, holland
from oracle.oracle import Oracle


def test_fort22_round_trip_and_columns():
    best = holland.synthetic_track(np.random.default_rng(1), -75.0, 27.0, 45.0, 6.0, 25.0, 950.0, 30.0)
    recs = holland.forecast_track(best, issue=3)
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "fort.22")
        holland.write_fort22(p, recs)
        np.testing.assert_array_equal(holland.read_fort22(p), recs)
    # a line in the ATCF layout the reference's FORMAT 14 expects (fields at fixed columns)
    line = ("AL, 12, 2005082800,   , BEST,   0, 257N,  870W, 145,  909, XX,  34, NEQ,  200,  200,"
            "  150,  150, 1008,  300,  15,   0,   0,   L,   0,    ,   0,   0,     KATRINA, D,")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "fort.22")
        open(p, "w").write(line + "\n")
        r = holland.read_fort22(p)[0]
    assert list(r) == [2005, 8, 28, 0, 0, 0, 257, 870, 145, 909, 300, 15]


def test_timeconv_is_day_of_year():
    for y, m, d, h in [(2020, 1, 1, 0), (2020, 3, 1, 6), (2021, 3, 1, 6), (2020, 12, 31, 18),
                       (2019, 8, 25, 12)]:
        doy = (datetime.datetime(y, m, d, h) - datetime.datetime(y, 1, 1)).total_seconds()
        assert holland.timeconv(y, m, d, h) == doy


def _oracle_for(case):
    return Oracle(case, threads=2)


def test_library_track_matches_oracle():
    case = load_case("storm3k")
    st = case.storm
    times = np.linspace(1.0, 40 * 3600.0, 97)
    lib_rows = gpu.holland_storm_at(st, case.params.g, times)
    o = _oracle_for(case)
    for t, row in zip(times, lib_rows):
        s = o.storm_at(t)
        # lat lon cpress deficit spd rmw tvx tvy B in the library's row order
        np.testing.assert_array_equal(row[:9], s)
        assert row[9] == math.cos(s[0] * (4.0 * math.atan(1.0) / 180.0))
        assert row[11] == st.bladj and row[12] == 1000.0 * case.params.g


def test_forecast_lines_use_nowcast_and_skip_repeats():
    """The advisory's repeated increment and its 0-h line are dropped; forecasts carry the last
    BEST line's pressure and radii; cast times follow the forecast increments."""
    case = load_case("storm3k")
    st = case.storm
    best = st.records[st.records[:, 4] == 0]
    now = best[-1]
    # times inside the hindcast-to-forecast interval: [-6 h, +6 h) around the reference time
    rows = gpu.holland_storm_at(st, case.params.g, np.array([-3 * 3600.0, 0.0, 5 * 3600.0]))
    assert np.all(rows[:, 2] == now[9] * 100.0)                       # central pressure (Pa)
    np.testing.assert_allclose(rows[:, 5], now[11] * 1.852000003180799)  # RMW (km)
    # halfway through the interval the position is the midpoint of nowcast and +12 h forecast
    ofcl = st.records[(st.records[:, 4] == 1) & (st.records[:, 5] == 12)][0]
    mid = gpu.holland_storm_at(st, case.params.g, np.array([0.0]))[0]
    assert mid[0] == pytest.approx(0.5 * (now[6] + ofcl[6]) / 10.0, abs=1e-12)
    assert mid[1] == pytest.approx(-0.5 * (now[7] + ofcl[7]) / 10.0, abs=1e-12)


def test_track_errors():
    case = load_case("storm3k")
    st = case.storm
    with pytest.raises(gpu.GpuError, match="does not cover"):
        gpu.holland_storm_at(st, 9.81, np.array([1.0e7]))
    bad = holland.HollandStorm(2020, 8, 25, 0, 1, 0.9, st.records[st.records[:, 4] == 1])
    with pytest.raises(gpu.GpuError, match="BEST"):
        gpu.holland_storm_at(bad, 9.81, np.array([0.0]))


def test_oracle_node_forcing_closed_form():
    """One stage's WSX2/WSY2/PR2 against the Holland profile written out independently:
    p(r) = pc + dp exp(-(Rmax/r)^B), Vg(r) from gradient balance, 10-m wind with the boundary-layer
    and 1-min-to-10-min factors plus the tapered translation, Garratt drag capped at 0.0035."""
    case = load_case("storm3k")
    o = _oracle_for(case)
    o.step(1)
    wsx, wsy, pr = o.met()
    p = case.params
    it, dt = 1, p.dtdp
    o2 = _oracle_for(case)
    s = o2.storm_at(it * dt + p.statim * 86400.0)
    lat, lon, pc, dp, spd, rmw, tvx, tvy, B = s
    nrk = case.dg.rk_stage
    # the last stage's ramp
    dtvd_last = 1.0      # SSP(2,2): DTVD = {0, 1}
    assert nrk == 2
    ramp = math.tanh((2.0 * ((it - 1) + dtvd_last) * dt / 86400.0) / p.dramp)
    R = 6378206.4
    omega = 2.0 * math.pi / 86164.2
    lam, phi = np.degrees(case.slam), np.degrees(case.sfea)
    dlam, dphi = np.radians(lam - lon), np.radians(phi - lat)
    a = np.sin(dphi / 2) ** 2 + math.cos(math.radians(lat)) * np.cos(np.radians(phi)) * np.sin(dlam / 2) ** 2
    r = 2 * R * np.arcsin(np.sqrt(a))
    f = 2 * omega * np.sin(np.radians(phi))
    x = (rmw * 1000.0 / r) ** B
    press = (pc + dp * np.exp(-x)) / (1000.0 * p.g)
    vg = np.sqrt(x * np.exp(1 - x) * spd ** 2 + (r * f / 2) ** 2) - r * f / 2
    theta = np.arctan2(dphi, dlam)
    wx = -vg * np.sin(theta) * st_bladj(case) * 0.88 + np.abs(vg) / spd * tvx
    wy = vg * np.cos(theta) * st_bladj(case) * 0.88 + np.abs(vg) / spd * tvy
    w = np.hypot(wx, wy)
    cd = np.minimum(0.001 * (0.75 + 0.067 * w), 0.0035)
    np.testing.assert_allclose(pr, ramp * press, rtol=1e-12)
    np.testing.assert_allclose(wsx, ramp * 0.001293 * cd * wx * w, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(wsy, ramp * 0.001293 * cd * wy * w, rtol=1e-9, atol=1e-15)
    # the profile peaks near Rmax: the strongest stress sits within a few Rmax of the eye
    k = int(np.argmax(np.hypot(wsx, wsy)))
    assert r[k] < 4 * rmw * 1000.0


def st_bladj(case):
    return case.storm.bladj
