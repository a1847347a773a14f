"""The fitted odd polynomials of the float halo profiles (atan_unit_f32 and
asinh_small_f32 in csrc/profiles.hpp), read from the source and evaluated in
float32 (one rounding per fused multiply-add) against numpy's atan / asinh over
their stated ranges.  A mistyped coefficient fails here, on the CPU."""

import os
import re

import numpy as np
import pytest

SRC = os.path.join(os.path.dirname(__file__), "..", "jaxtronomy_amd", "csrc", "profiles.hpp")


def coefficients(name):
    text = open(SRC).read()
    body = re.search(r"float %s\(float \w\) \{(.*?)\n\}" % name, text, re.S).group(1)
    nums = [float(x) for x in re.findall(r"(-?\d+\.\d+(?:e-?\d+)?)f", body)]
    # highest power first: the initial value, then one fmaf constant per step
    return nums


def eval_f32(coefs, t):
    f32 = np.float32
    t = f32(t)
    s = f32(t * t)
    p = f32(coefs[0])
    for c in coefs[1:]:
        p = f32(np.float64(p) * np.float64(s) + np.float64(f32(c)))
    ps = f32(p * s)
    return f32(np.float64(t) * np.float64(ps) + np.float64(t))


@pytest.mark.parametrize("name,fn,hi,tol_ulp", [("atan_unit_f32", np.arctan, 0.9, 2.0),
                                                ("asinh_small_f32", np.arcsinh, 0.5, 1.5)])
def test_fitted_polynomial_accuracy(name, fn, hi, tol_ulp):
    coefs = coefficients(name)
    assert len(coefs) >= 5
    t = np.concatenate([np.geomspace(1e-6, hi, 3000), np.linspace(0.0, hi, 3001)[1:]])
    got = np.array([eval_f32(coefs, x) for x in t], dtype=np.float64)
    ref = fn(t.astype(np.float32).astype(np.float64))
    rel = np.max(np.abs(got - ref) / ref)
    assert rel < tol_ulp * 2.0 ** -24, (name, rel)


def test_far_field_ranges_fit():
    """The callers stay inside the fitted ranges: 1/sqrt(X^2 - 1) <= 0.9 for
    X >= 1.5 (TNFW and NFW far field), and the asinh series is used for
    tau / X <= 0.5 only."""
    assert 1.0 / np.sqrt(1.5 ** 2 - 1.0) <= 0.9
    text = open(SRC).read()
    assert "X >= T(1.5)" in text and "X >= 1.5f" in text
    assert "float ash = asinh_small_f32(z);" in text and "if (z > 0.5f) ash = " in text
