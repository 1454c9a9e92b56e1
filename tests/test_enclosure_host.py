"""CPU property test of the mixed-precision AABB enclosure: for polynomials of
every compiled order (2-20) with adversarial coefficients and intervals, the
FP32 Cargo-Shisha enclosure with its a-priori margin contains the FP64
enclosure the FP64 screen computes, and the exact range of the polynomial at
sample points (rational arithmetic on the FP64 coefficients). Whenever the FP32
result is not finite the kernel uses the FP64 enclosure instead, so such cases
only count. The routines are the kernels' own (mzh_enclosure.hpp), built for
the host with the same -ffp-contract=off."""

import ctypes
import pathlib
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def enc(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    out = tmp_path_factory.mktemp("enc") / "libenc.so"
    subprocess.run(
        [gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-fPIC", "-shared",
         str(ROOT / "tests" / "enclosure_check.cpp"), "-o", str(out)],
        check=True,
    )
    lib = ctypes.CDLL(str(out))
    lib.enc_batch.restype = ctypes.c_int
    return lib


def _run(lib, order, cf, lb, ub):
    n = len(lb)
    cf = np.ascontiguousarray(cf, dtype=np.float64)
    lb = np.ascontiguousarray(lb, dtype=np.float64)
    ub = np.ascontiguousarray(ub, dtype=np.float64)
    e64 = np.zeros((n, 2))
    e32 = np.zeros((n, 2), dtype=np.float32)
    P = ctypes.POINTER(ctypes.c_double)
    rc = lib.enc_batch(order, ctypes.c_uint64(n), cf.ctypes.data_as(P), lb.ctypes.data_as(P), ub.ctypes.data_as(P),
                       e64.ctypes.data_as(P), e32.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert rc == 0
    return e64, e32


def _cases(order, n, rng):
    """Coefficient families: random magnitudes over 40 decades with random
    signs; near-cancelling products (t - r)^k with clustered roots; position-
    like expansions of 7000 cos(w t + phi) (large constant term, alternating
    decreasing terms); and a huge leading term. Intervals from 1e-12 to 1e2
    long, starting anywhere in [0, 3)."""
    cfs, lbs, ubs = [], [], []
    for k in range(n):
        fam = k % 4
        if fam == 0:
            c = rng.choice([-1.0, 1.0], order + 1) * 10.0 ** rng.uniform(-20, 20, order + 1)
        elif fam == 1:
            r = rng.uniform(0, 3)
            c = np.array([1.0])
            for _ in range(order):
                c = np.convolve(c, [1.0, -(r + rng.normal(0, 10.0 ** rng.uniform(-9, -1)))])
            c = c[::-1] * 10.0 ** rng.uniform(-3, 3)
        elif fam == 2:
            w, ph, a = rng.uniform(0.5, 80.0), rng.uniform(0, 2 * np.pi), rng.uniform(6500, 42000)
            c = np.array([a * np.cos(ph + np.pi * m / 2) * w**m / np.prod(np.arange(1, m + 1)) for m in range(order + 1)])
        else:
            c = rng.uniform(-1, 1, order + 1)
            c[-1] = rng.choice([-1.0, 1.0]) * 10.0 ** rng.uniform(10, 30)
        lb = rng.uniform(0, 3) if rng.random() < 0.7 else 0.0
        ub = lb + 10.0 ** rng.uniform(-12, 2)
        cfs.append(c)
        lbs.append(lb)
        ubs.append(ub)
    return np.array(cfs), np.array(lbs), np.array(ubs)


def _exact(c, t):
    v = Fraction(0)
    for x in c[::-1]:
        v = v * t + Fraction(float(x))
    return v


@pytest.mark.parametrize("order", list(range(2, 21)))
def test_fp32_enclosure_contains_fp64(enc, order):
    rng = np.random.default_rng(100 + order)
    cf, lb, ub = _cases(order, 400, rng)
    e64, e32 = _run(enc, order, cf, lb, ub)
    fin = np.isfinite(e32).all(axis=1) & np.isfinite(e64).all(axis=1)
    assert fin.sum() >= 250
    lo32 = e32[:, 0].astype(np.float64)
    hi32 = e32[:, 1].astype(np.float64)
    # the FP64 enclosure inside the FP32 one (float -> double is exact)
    assert np.all(lo32[fin] <= e64[fin, 0]), np.flatnonzero(fin & (lo32 > e64[:, 0]))
    assert np.all(hi32[fin] >= e64[fin, 1]), np.flatnonzero(fin & (hi32 < e64[:, 1]))
    # the exact range at the ends and inside
    for k in np.flatnonzero(fin)[::4]:
        a, b = Fraction(float(lb[k])), Fraction(float(ub[k]))
        for s in (Fraction(0), Fraction(1, 3), Fraction(1, 2), Fraction(1)):
            v = _exact(cf[k], a + (b - a) * s)
            assert Fraction(float(lo32[k])) <= v <= Fraction(float(hi32[k])), (k, float(s))


def test_fp32_margin_is_tight_for_orbits(enc):
    """The margin is small where it matters: for position-like polynomials
    (km-sized constant term, alternating decreasing terms) the FP32 interval
    exceeds the FP64 one by a median of under 1e-4 of the enclosure's
    magnitude (under a kilometre at 7000 km), so the mixed screen's boxes, and
    its pair count, stay close to the FP64 screen's."""
    for order in range(2, 13):
        rng = np.random.default_rng(200 + order)
        cf, lb, ub = _cases(order, 400, rng)
        e64, e32 = _run(enc, order, cf, lb, ub)
        m = (np.arange(400) % 4 == 2) & np.isfinite(e32).all(axis=1)
        slack = (e64[m, 0] - e32[m, 0].astype(np.float64)) + (e32[m, 1].astype(np.float64) - e64[m, 1])
        scale = np.maximum(np.abs(e64[m, 0]), np.abs(e64[m, 1]))
        assert np.all(slack >= 0)
        assert np.median(slack / scale) < 1e-4, order
