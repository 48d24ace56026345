"""Accuracy and semantics of the shared elementary functions (include/wfh_math.h).

The kernels and the oracle evaluate exp, log, pow, cbrt, cld, round(sigdigits=12), min and max
with the same code, so they agree bit for bit; these tests pin that code against independent
references: numpy/libm to within a stated number of ulps on random samples over the full
range, exact rational arithmetic for cld, decimal rounding for round(sigdigits=12), and
Julia's documented min/max semantics for NaN and signed zeros."""
from fractions import Fraction
import math

import numpy as np

import oracle as O


def ulps(a, b):
    """Distance in units of the last place of b."""
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_exp_within_one_ulp():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-708.0, 709.0, 400_000), rng.uniform(-1.0, 1.0, 200_000),
                        rng.uniform(-1e-8, 1e-8, 50_000), np.array([0.0, -0.0, 1.0, -1.0, 709.78, -708.39])])
    got = O.math("exp", x)
    assert ulps(got, np.exp(x)).max() <= 1.0


def test_exp_subnormal_and_special():
    x = np.array([-740.0, -744.0, -745.2, 710.0, np.inf, -np.inf, np.nan])
    got = O.math("exp", x)
    with np.errstate(over="ignore"):
        ref = np.exp(x)
    fin = np.isfinite(ref) & (ref > 0)
    # subnormal results: within one subnormal step of libm
    assert np.all(np.abs(got[fin] - ref[fin]) <= np.spacing(0.0) * 1.0 + np.spacing(ref[fin]))
    assert got[2] == 0.0 and got[3] == np.inf and got[4] == np.inf and got[5] == 0.0 and np.isnan(got[6])


def test_log_within_one_ulp():
    rng = np.random.default_rng(2)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 400_000), rng.uniform(0.5, 2.0, 200_000),
                        1.0 + rng.uniform(-1e-6, 1e-6, 50_000), np.array([1e-310, 5e-324, 2.0, 0.5, 1.0])])
    got = O.math("log", x)
    ref = np.log(x)
    nz = ref != 0.0
    assert ulps(got[nz], ref[nz]).max() <= 1.0
    assert got[~nz].tolist() == [0.0] * int((~nz).sum())


def test_log_special():
    got = O.math("log", np.array([0.0, -1.0, np.inf, np.nan]))
    assert got[0] == -np.inf and np.isnan(got[1]) and got[2] == np.inf and np.isnan(got[3])


def test_pow_is_exp_of_y_log_x():
    """pow(x, y) = exp(y log x) (the reference's own composition, Wflow/src/utils.jl:470):
    a one-ulp error of log(x) becomes an absolute error of about |y log x| ulp-units in the
    argument of exp, i.e. a relative error of the result, so the bound is 2 + 2 |y log x| ulps."""
    rng = np.random.default_rng(3)
    x = 10.0 ** rng.uniform(-30, 10, 300_000)
    y = rng.choice([0.2, 0.6, 1.0 / 3.0, 2.5, -0.4, 1.7], x.size)
    got = O.math("pow", x, y)
    ref = np.power(x, y)
    bound = 2.0 + 2.0 * np.abs(y * np.log(x))
    assert np.all(ulps(got, ref) <= bound)


def test_cbrt_within_one_ulp():
    rng = np.random.default_rng(4)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 200_000), -(10.0 ** rng.uniform(-30, 30, 50_000)),
                        np.array([8.0, 27.0, 1.0, -8.0])])
    got = O.math("cbrt", x)
    assert ulps(got, np.cbrt(x)).max() <= 1.0
    assert got[-4:].tolist() == [2.0, 3.0, 1.0, -2.0]


def test_cld_exact():
    """cld(x, y) = ceil of the exact quotient (Julia's float cld), Wflow soil_process.jl:74."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0.0, 0.05, 20_000), 2e-4 * np.arange(1, 2000),
                        np.nextafter(2e-4 * np.arange(1, 200), 0.0), np.nextafter(2e-4 * np.arange(1, 200), 1.0)])
    got = O.math("cld", x, 2e-4)
    q = Fraction(2e-4)
    ref = np.array([math.ceil(Fraction(v) / q) for v in x], dtype=np.float64)
    assert np.array_equal(got, ref)


def julia_round_sig12(v):
    """Julia's round(v; sigdigits=12) restated with exact integer arithmetic for the decimal
    exponent and the correctly rounded power of ten (Base._round_sigdigits)."""
    a = Fraction(abs(v))
    e = len(str(a.numerator // a.denominator)) - 1 if a >= 1 else -len(str(a.denominator // a.numerator))
    while Fraction(10) ** e > a:
        e -= 1
    while Fraction(10) ** (e + 1) <= a:
        e += 1
    d = 12 - (1 + e)
    if d >= 0:
        sc = float(10 ** d)
        return float(round(v * sc)) / sc
    isc = float(10 ** -d)
    return float(round(v / isc)) * isc


def test_pow10_correctly_rounded():
    """10.0^n as Julia's compensated power by squaring: the correctly rounded power."""
    n = np.arange(0, 309, dtype=np.float64)
    got = O.math("pow10", n)
    assert np.array_equal(got, np.array([float(10 ** int(k)) for k in n]))


def test_round_sig12_matches_julia_algorithm():
    rng = np.random.default_rng(6)
    # the range where every power of ten involved is exact (see wfh_round_sig12)
    x = np.concatenate([10.0 ** rng.uniform(-10, 10, 30_000), rng.uniform(0.5, 30.0, 30_000),
                        -rng.uniform(0.5, 30.0, 1000), np.array([0.1 * 3, 1.0, 12.5, 0.99949999999999994])])
    got = O.math("round_sig12", x)
    ref = np.array([julia_round_sig12(float(v)) for v in x])
    assert np.array_equal(got, ref)
    # and it is the 12-significant-digit decimal rounding wherever the decimal value is not a tie
    mid = (x > 1e-3) & (x < 1e3)
    dec = np.array([float(f"{v:.11e}") for v in x[mid]])
    assert np.mean(got[mid] == dec) > 0.999


def test_min_max_julia_semantics():
    nan = float("nan")
    x = np.array([nan, 1.0, -0.0, 0.0, 2.0, -np.inf])
    y = np.array([1.0, nan, 0.0, -0.0, 3.0, 5.0])
    mn = O.math("min", x, y)
    mx = O.math("max", x, y)
    assert np.isnan(mn[0]) and np.isnan(mn[1]) and np.isnan(mx[0]) and np.isnan(mx[1])
    assert np.signbit(mn[2]) and np.signbit(mn[3]) and not np.signbit(mx[2]) and not np.signbit(mx[3])
    assert mn[4] == 2.0 and mx[4] == 3.0 and mn[5] == -np.inf and mx[5] == 5.0
