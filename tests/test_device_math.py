"""CPU pins of the float32 algorithms the kernels run (emulated in numpy).

fast_sincos (emsim_amd/csrc/fft_lds.hpp): Cody-Waite reduction by pi/2 with
a three-part split and Taylor polynomials.  The emulation follows the device
code step by step; FMA is emulated by one float64 operation rounded to
float32 (exact for these operands up to a double rounding).
"""
import math
import re
from pathlib import Path

import numpy as np
import pytest

F = np.float32
SRC = Path(__file__).resolve().parent.parent / "emsim_amd" / "csrc" / "fft_lds.hpp"


def _constants():
    text = SRC.read_text()
    body = text[text.index("fast_sincos"):]
    nums = [float(x.rstrip("f")) for x in re.findall(r"[-+]?\d\.\d+(?:e[-+]\d+)?f", body)]
    return nums


def fma(a, b, c):
    return F(np.float64(a) * np.float64(b) + np.float64(c))


def fast_sincos(x):
    c = _constants()
    two_over_pi, p1, p2, p3 = (F(v) for v in c[0:4])
    sin_coef, cos_coef = c[4:9], c[9:15]
    x = F(x)
    kf = F(np.rint(F(x * two_over_pi)))
    r = fma(-kf, p1, x)
    r = fma(-kf, p2, r)
    r = fma(-kf, p3, r)
    r2 = F(r * r)
    ps = F(sin_coef[0])
    for v in sin_coef[1:]:
        ps = fma(ps, r2, F(v))
    sn = fma(F(ps * r2), r, r)
    pc = F(cos_coef[0])
    for v in cos_coef[1:]:
        pc = fma(pc, r2, F(v))
    cs = fma(pc, r2, F(1.0))
    q = int(kf) & 3
    s, co = (cs, sn) if q & 1 else (sn, cs)
    if q in (1, 2):
        co = -co
    if q & 2:
        s = -s
    return s, co


def test_constants_are_the_intended_ones():
    c = _constants()
    assert "fabsf(x) <= 8192.f" in SRC.read_text()
    assert abs(c[0] - 2 / math.pi) < 1e-16
    assert abs(c[1] + c[2] + c[3] - math.pi / 2) < 1e-14
    # Taylor coefficients of sin (x^11 .. x^3) and cos (x^12 .. x^2)
    for v, k in zip(c[4:9], (11, 9, 7, 5, 3)):
        assert abs(abs(v) * math.factorial(k) - 1) < 1e-12
    for v, k in zip(c[9:15], (12, 10, 8, 6, 4, 2)):
        assert abs(abs(v) * math.factorial(k) - 1) < 1e-12


@pytest.mark.parametrize("lo,hi", [(-8192, 8192), (-10, 10), (0, 0.1)])
def test_fast_sincos_accuracy(lo, hi):
    xs = np.random.default_rng(int(hi)).uniform(lo, hi, 4000).astype(F)
    xs = np.concatenate([xs, (np.arange(-16, 17) * np.pi / 4).astype(F), np.array([0.0, -0.0], F)])
    err = 0.0
    for x in xs:
        s, c = fast_sincos(x)
        err = max(err, abs(float(s) - math.sin(float(x))), abs(float(c) - math.cos(float(x))))
    assert err < 1.2e-7   # about one float32 ulp of 1
