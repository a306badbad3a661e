"""The float32 DSF pair's erfc (csrc/pairkern.h, erfc_from_exp): erfc(x) = exp(-x^2) R(t) with
t = 1 / (1 + p x), R a degree-8 polynomial in t without constant term (the form of Abramowitz
and Stegun eq. 7.1.26, coefficients fitted for relative error on [0, 4]). The constants are read
from the source, evaluated in float32 in the kernel's operation order, and checked against
scipy's erfc: relative error below 1e-6 on [0, 4] and below 7e-7 on [0, 2.4] (the 65,536-ion
NaCl workload's alpha rc), and R(1) = erfc(0) = 1 to float32 rounding."""
import os
import re

import numpy as np
import pytest

scipy_special = pytest.importorskip("scipy.special")

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "devinproj_amd", "csrc", "pairkern.h")


def _constants():
    text = open(SRC).read()
    body = text[text.index("float erfc_from_exp("):]
    body = body[:body.index("\n}\n")]
    p = float(re.search(r"fmaf\(([-0-9.e]+)f, x, 1\.0f\)", body).group(1))
    c_top = float(re.search(r"float r = ([-0-9.e]+)f;", body).group(1))
    rest = [float(v) for v in re.findall(r"r = fmaf\(r, t, ([-0-9.e]+)f\);", body)]
    return p, [c_top] + rest  # highest degree first


def _erfc32(x):
    p, coeffs = _constants()
    x = x.astype(np.float32)
    t = (np.float32(1) / (np.float32(p) * x + np.float32(1))).astype(np.float32)
    r = np.full_like(x, np.float32(coeffs[0]))
    for c in coeffs[1:]:
        r = (r * t + np.float32(c)).astype(np.float32)
    ex = np.exp(-(x * x)).astype(np.float32)
    return (ex * (r * t).astype(np.float32)).astype(np.float32)


def test_constants_are_a_degree_8_polynomial():
    p, coeffs = _constants()
    assert len(coeffs) == 8 and 0.3 < p < 0.5
    assert abs(sum(coeffs) - 1.0) < 1e-6  # t = 1 at x = 0: erfc(0) = 1


def test_relative_error_against_erfc():
    x = np.linspace(0.0, 4.0, 400001)
    ref = scipy_special.erfc(x.astype(np.float32).astype(np.float64))
    rel = np.abs(_erfc32(x).astype(np.float64) / ref - 1.0)
    assert rel.max() < 1e-6, (rel.max(), x[rel.argmax()])
    assert rel[x <= 2.4].max() < 7e-7
