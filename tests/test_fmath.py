"""Accuracy of the shared elementary functions (mray_amd/csrc/fmath.h) that
both the gfx950 kernels and the oracle use, against numpy in double over the
ranges the path evaluates them on (and beyond): error in units in the last
place of the float result, plus the special values."""
import math

import numpy as np
import pytest

from oracle import refcpu


def _ulp_err(got, want):
    want32 = want.astype(np.float32)
    ulp = np.spacing(np.abs(want32)).astype(np.float64)
    ulp = np.maximum(ulp, np.float64(np.spacing(np.float32(0))))
    return np.abs(got.astype(np.float64) - want) / ulp


def _grid(a, b, n=400_001):
    return np.linspace(a, b, n).astype(np.float32)


CASES = [
    # name, argument range, numpy reference, max ulp (measured + margin)
    ("sin", (-4 * math.pi, 4 * math.pi), np.sin, None),
    ("cos", (-4 * math.pi, 4 * math.pi), np.cos, None),
    ("exp", (-87.0, 88.0), np.exp, 2.0),
    ("log", (1e-30, 1e30), np.log, 2.0),
    ("acos", (-1.0, 1.0), np.arccos, 3.0),
    ("asin", (-1.0, 1.0), np.arcsin, None),
    ("atanh", (-0.85691062, 0.97059135), np.arctanh, None),
    ("cosh", (-3.5, 3.5), np.cosh, 3.0),
]


@pytest.mark.parametrize("name,rng,ref,bound", CASES)
def test_fmath_accuracy(name, rng, ref, bound):
    if name == "log":
        x = np.exp(np.linspace(np.log(rng[0]), np.log(rng[1]), 400_001)).astype(np.float32)
    else:
        x = _grid(*rng)
    got = refcpu.fmath(name, x)
    want = ref(x.astype(np.float64))
    if bound is not None:
        assert _ulp_err(got, want).max() <= bound, (name, _ulp_err(got, want).max())
    else:
        # functions with zeros inside the range: absolute error against the
        # result's scale (ulp of 1 for |result| <= 1), and relative away from zeros
        abs_err = np.abs(got.astype(np.float64) - want)
        assert abs_err.max() <= 2.0 * np.spacing(np.float32(1.0)), (name, abs_err.max())
        big = np.abs(want) > 1e-3
        assert _ulp_err(got[big], want[big]).max() <= 4.0, (name, _ulp_err(got[big], want[big]).max())


def test_atan2_all_quadrants():
    rng = np.random.default_rng(3)
    y = rng.normal(size=300_000).astype(np.float32)
    x = rng.normal(size=300_000).astype(np.float32)
    got = refcpu.fmath("atan2", y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert _ulp_err(got, want).max() <= 3.0
    # axes, zeros of either sign, infinities
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, 0.0, np.inf, -np.inf, np.inf], np.float32)
    sx = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, -np.inf, -np.inf, np.inf, 1.0, -np.inf],
                  np.float32)
    assert np.array_equal(refcpu.fmath("atan2", sy, sx),
                          np.arctan2(sy, sx).astype(np.float32))


def test_special_values():
    assert np.isnan(refcpu.fmath("log", np.array([-1.0]))[0])
    assert refcpu.fmath("log", np.array([0.0]))[0] == -np.inf
    assert refcpu.fmath("log", np.array([1.0]))[0] == 0.0
    assert refcpu.fmath("exp", np.array([0.0]))[0] == 1.0
    assert refcpu.fmath("exp", np.array([100.0]))[0] == np.inf
    assert refcpu.fmath("exp", np.array([-110.0]))[0] == 0.0
    sub = refcpu.fmath("exp", np.array([-100.0]))[0]
    assert sub > 0 and abs(sub / math.exp(-100.0) - 1) < 1e-2      # subnormal result
    assert refcpu.fmath("log", np.array([1e-40], np.float32))[0] == pytest.approx(
        math.log(np.float32(1e-40)), rel=1e-6)                        # subnormal argument
    assert refcpu.fmath("sin", np.array([0.0]))[0] == 0.0
    assert refcpu.fmath("cos", np.array([0.0]))[0] == 1.0
    assert refcpu.fmath("acos", np.array([1.0]))[0] == 0.0
    assert refcpu.fmath("acos", np.array([-1.0]))[0] == np.float32(math.pi)
    assert np.isnan(refcpu.fmath("sin", np.array([np.inf]))[0])


def test_sampling_ranges_exactly_as_used():
    """The path's arguments: 2 pi xi for xi = k 2^-32 (cosine and VNDF
    sampling), and the hyperbolic wavelength sampler's atanh / cosh."""
    xi = (np.arange(0, 1 << 32, 1 << 12, dtype=np.float64) * 2.0 ** -32).astype(np.float32)
    ang = (np.float32(2.0 * math.pi) * xi).astype(np.float32)
    s = refcpu.fmath("sin", ang)
    c = refcpu.fmath("cos", ang)
    a64 = ang.astype(np.float64)
    assert np.abs(s - np.sin(a64)).max() <= 1.5 * np.spacing(np.float32(1.0))
    assert np.abs(c - np.cos(a64)).max() <= 1.5 * np.spacing(np.float32(1.0))
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 3e-7
