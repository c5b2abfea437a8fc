"""The step's correctly rounded cos and x**(1/3) (dgswem_amd/csrc/dg_crmath.hpp), built for the host.

The GPU kernels evaluate the reference's COS (tidal boundary, src/ocean_edge_hydro.F:459-488) and
DEPTH**(1.D0/3.D0) (Manning, src/rhs_dg_hydro.F:302) with these functions.  Checked here against
the exact values: each seeded argument's result must equal the double nearest to the true value,
computed with 90-digit decimal arithmetic (Taylor series for cos, exp/ln for the power with the
exponent 1/3 rounded to a double).  Special arguments (exact cubes, zero, negative and
out-of-range inputs, arguments near multiples of pi/2) are checked against their closed forms.
The C library's own cos and pow are not always correctly rounded; the test also records how
often they differ (a fraction of a percent) to document that the functions stand in for them.
"""
import ctypes
import math
import os
import random
import subprocess
from decimal import Decimal, getcontext
from fractions import Fraction

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "dgswem_amd", "csrc", "dg_crmath.hpp")

SRC = r"""
#include "%s"
extern "C" double t_cos(double x) { return dg::crm::cr_cos(x); }
extern "C" double t_pow3(double x) { return dg::crm::cr_pow_third(x); }
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("crm")
    src = d / "t.cpp"
    src.write_text(SRC % os.path.abspath(HEADER))
    so = d / "libt.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-o", str(so), str(src)],
                   check=True)
    L = ctypes.CDLL(str(so))
    for f in (L.t_cos, L.t_pow3):
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_double]
    return L


getcontext().prec = 90


def _dec(x: float) -> Decimal:
    f = Fraction(x)
    return Decimal(f.numerator) / Decimal(f.denominator)


def _nearest(v: Decimal) -> float:
    f = float(v)
    cands = [f, math.nextafter(f, math.inf), math.nextafter(f, -math.inf)]
    return min(cands, key=lambda c: abs(_dec(c) - v))


def exact_cos(x: float) -> float:
    xd = _dec(x)
    s, t, k = Decimal(0), Decimal(1), 0
    while abs(t) > Decimal(10) ** -85:
        s += t
        k += 2
        t = -t * xd * xd / (k * (k - 1))
    return _nearest(s)


THIRD = Fraction(1.0 / 3.0)


def exact_pow3(x: float) -> float:
    y = Decimal(THIRD.numerator) / Decimal(THIRD.denominator)
    return _nearest((_dec(x).ln() * y).exp())


def test_cos_is_correctly_rounded(lib):
    rng = random.Random(7)
    args = [rng.uniform(-20.0, 40.0) for _ in range(1500)] + [rng.uniform(-1e-3, 1e-3) for _ in range(100)]
    bad = [(x, lib.t_cos(x), exact_cos(x)) for x in args if lib.t_cos(x) != exact_cos(x)]
    assert not bad, bad[:5]


def test_cos_near_multiples_of_half_pi(lib):
    for k in range(-12, 25):
        x = k * math.pi / 2
        for y in (x, math.nextafter(x, math.inf), math.nextafter(x, -math.inf)):
            assert lib.t_cos(y) == exact_cos(y), (k, y)
    assert lib.t_cos(0.0) == 1.0
    assert math.isnan(lib.t_cos(float("nan")))


def _pi_decimal() -> Decimal:
    """pi to about 110 digits (Machin's formula in exact rationals)."""
    def atan_inv(n, terms):
        s, x = Fraction(0), Fraction(1, n)
        t = x
        for k in range(terms):
            s += (t / (2 * k + 1)) * (-1) ** k
            t *= x * x
        return s
    pi = 16 * atan_inv(5, 160) - 4 * atan_inv(239, 50)
    return Decimal(pi.numerator) / Decimal(pi.denominator)


def exact_cos_reduced(x: float) -> float:
    """cos(x) for large |x|: reduced by pi/2 in 90-digit arithmetic first, then the series."""
    xd, hp = _dec(x), _pi_decimal() / 2
    k = int((xd / hp).to_integral_value())
    r = xd - k * hp
    cs, sn, t, j = Decimal(0), Decimal(0), Decimal(1), 0
    while abs(t) > Decimal(10) ** -88:
        if j % 2 == 0:
            cs += t if j % 4 == 0 else -t
        else:
            sn += t if j % 4 == 1 else -t
        j += 1
        t = t * r / j
    v = [cs, -sn, -cs, sn][k % 4]
    return _nearest(v)


def test_cos_near_multiples_of_half_pi_at_the_largest_accepted_arguments(lib):
    """The reduction uses pi/2 in four parts with k*P3 an exact product: arguments next to
    k*pi/2 keep their relative accuracy up to the largest k accepted (|x| < 2^16, k < 41722)."""
    hp = _pi_decimal() / 2
    for k in (41721, 40000, 33333, 2 ** 15, 12345, 1001):
        x = float(k * hp)
        assert abs(x) < 65536.0
        for y in (x, math.nextafter(x, math.inf), math.nextafter(x, -math.inf), -x):
            assert lib.t_cos(y) == exact_cos_reduced(y), (k, y)
    rng = random.Random(5)
    for _ in range(200):
        y = rng.uniform(1000.0, 65535.0)
        assert lib.t_cos(y) == exact_cos_reduced(y), y


def test_pow_third_is_correctly_rounded(lib):
    rng = random.Random(11)
    args = [math.exp(rng.uniform(-9.0, 9.5)) for _ in range(1500)]
    bad = [(x, lib.t_pow3(x), exact_pow3(x)) for x in args if lib.t_pow3(x) != exact_pow3(x)]
    assert not bad, bad[:5]


def test_pow_third_special_arguments(lib):
    # exact cubes: x**(1/3 - 2^-54/3) sits just below the cube root c, and rounds to c or to the
    # double below it (1000**y is 9.999999999999998, as the C library's pow also returns)
    for c in (0.5, 1.0, 2.0, 3.0, 10.0, 1234.0):
        assert lib.t_pow3(c ** 3) == exact_pow3(c ** 3)
    assert lib.t_pow3(8.0) == 2.0 and lib.t_pow3(1000.0) == 9.999999999999998
    assert lib.t_pow3(0.0) == 0.0 and math.copysign(1.0, lib.t_pow3(-0.0)) == 1.0
    assert math.isnan(lib.t_pow3(-8.0))
    assert lib.t_pow3(1e-310) == pytest.approx(1e-310 ** (1.0 / 3.0), rel=1e-15)


def test_close_to_the_c_library(lib):
    """The C library rounds a small fraction of these arguments the other way."""
    libm = ctypes.CDLL("libm.so.6")
    libm.cos.restype = libm.pow.restype = ctypes.c_double
    libm.cos.argtypes = [ctypes.c_double]
    libm.pow.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = random.Random(3)
    n = 200_000
    dc = sum(lib.t_cos(x) != libm.cos(x) for x in (rng.uniform(-20.0, 40.0) for _ in range(n)))
    dp = sum(lib.t_pow3(x) != libm.pow(x, 1.0 / 3.0) for x in (math.exp(rng.uniform(-9.0, 9.5)) for _ in range(n)))
    print(f"differ from the C library: cos {dc / n:.2e}, pow {dp / n:.2e}")
    assert dc / n < 5e-3 and dp / n < 5e-3
