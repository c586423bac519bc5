"""Pins of the oracle's numerical building blocks: deterministic pow/log against mpmath,
Philox4x32-10 against the Random123 known-answer vectors, the seeded permeability field
against its distribution, and the product's host-compiled copies bit for bit against the
oracle's."""
import math

import mpmath
import numpy as np
import pytest

from oracle import oracle_ffi as O
from tests.helpers import ensure_oracle, hostcheck

mpmath.mp.prec = 160


@pytest.fixture(scope="module", autouse=True)
def _lib():
    ensure_oracle()


def _ulps(got, x, y):
    ex = mpmath.power(mpmath.mpf(x), mpmath.mpf(y))
    return abs(float((mpmath.mpf(got) - ex) / math.ulp(float(ex))))


@pytest.mark.parametrize("lo,hi,ylo,yhi,bound", [
    (1e-3, 1.0, -3.0, 3.0, 0.75),        # van Genuchten arguments (swe^(1/m), (x-1)^(1/n), ...)
    (0.5, 2.0, -3.0, 3.0, 0.75),
    (1e-300, 1e300, -0.9, 0.9, 0.75),    # wide exponents
])
def test_dpow_accuracy(lo, hi, ylo, yhi, bound):
    rng = np.random.default_rng(7)
    xs = np.exp(rng.uniform(np.log(lo), np.log(hi), 3000))
    ys = rng.uniform(ylo, yhi, 3000)
    worst = max(_ulps(O.dpow(x, y), x, y) for x, y in zip(xs, ys))
    assert worst < bound, worst


def test_dpow_pow10_range():
    rng = np.random.default_rng(3)
    zs = rng.uniform(-16.0, -6.0, 3000)
    worst = max(_ulps(O.dpow(10.0, z), 10.0, z) for z in zs)
    assert worst < 0.75


def test_dpow_special_cases():
    assert O.dpow(2.0, 0.0) == 1.0
    assert O.dpow(1.0, 123.0) == 1.0
    assert O.dpow(0.0, 0.5) == 0.0
    assert O.dpow(0.0, -0.5) == math.inf
    assert O.dpow(4.0, 0.5) == 2.0
    assert O.dpow(2.0, 10.0) == 1024.0
    assert O.dpow(0.25, -0.5) == 2.0
    assert math.isnan(O.dpow(-1.0, 0.5))
    assert O.dpow(5e-324, 1.0) == 5e-324  # subnormal input


def test_dlog_accuracy():
    rng = np.random.default_rng(5)
    xs = np.exp(rng.uniform(-700, 700, 3000))
    for x in xs:
        got = O.dlog(x)
        ex = mpmath.log(mpmath.mpf(x))
        assert abs(float((mpmath.mpf(got) - ex) / math.ulp(float(ex)))) < 0.75


def test_product_pow_bitwise_equals_oracle():
    H = hostcheck()
    rng = np.random.default_rng(11)
    xs = np.exp(rng.uniform(-30, 30, 20000))
    ys = rng.uniform(-4, 4, 20000)
    for x, y in zip(xs, ys):
        a, b = O.dpow(x, y), H.hc_dpow(x, y)
        assert np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64), (x, y)
    for x in xs[:2000]:
        assert O.dlog(x) == H.hc_dlog(x)


# Random123 kat_vectors for philox4x32_10
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert O.philox(ctr, key) == out
    import ctypes as C
    o = (C.c_uint32 * 4)()
    hostcheck().hc_philox((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), o)
    assert list(o) == out


def test_lognormal_distribution():
    K = O.lognormal_K(42, 7, 200000, -11.5, 0.75)
    z = np.log10(K)
    assert abs(z.mean() + 11.5) < 0.01
    assert abs(z.std() - 0.75) < 0.01
    from scipy import stats
    assert stats.kstest((z + 11.5) / 0.75, "norm").pvalue > 1e-3
    # different problem ids / seeds give different fields
    assert not np.array_equal(K[:100], O.lognormal_K(42, 8, 100, -11.5, 0.75))
