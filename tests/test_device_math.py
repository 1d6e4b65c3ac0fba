"""CPU pins of the float32 algorithms the kernels run (emulated in numpy).

fast_sincos (emsim_amd/csrc/fft_lds.hpp): Cody-Waite reduction by pi/2 with
a three-part split and fitted polynomials (sin r = r + r^3 p(r^2), cos r =
1 - r^2/2 + r^4 q(r^2)).  The emulation follows the device code step by
step; FMA is emulated by one float64 operation rounded to float32 (exact for
these operands up to a double rounding).
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
    x = F(x)
    kf = F(np.rint(F(x * two_over_pi)))
    r = fma(-kf, p1, x)
    r = fma(-kf, p2, r)
    r = fma(-kf, p3, r)
    sn, cs = _poly(r)
    q = int(kf) & 3
    s, co = (cs, sn) if q & 1 else (sn, cs)
    if q in (1, 2):
        co = -co
    if q & 2:
        s = -s
    return s, co


def _poly(r):
    """sin r and cos r on [-pi/4, pi/4], as the device evaluates them."""
    c = _constants()
    s3, s2, s1, c3, c2, c1, half, one = (F(v) for v in c[4:12])
    r2 = F(r * r)
    ps = fma(fma(s3, r2, s2), r2, s1)
    sn = fma(ps, F(r * r2), r)
    pc = fma(fma(c3, r2, c2), r2, c1)
    cs = fma(pc, F(r2 * r2), fma(half, r2, one))
    return sn, cs


def test_constants_are_the_intended_ones():
    c = _constants()
    assert "fabsf(x) <= 8192.f" in SRC.read_text()
    assert abs(c[0] - 2 / math.pi) < 1e-16
    assert abs(c[1] + c[2] + c[3] - math.pi / 2) < 1e-14
    # the fitted coefficients stay close to the Taylor ones they refine
    # (sin: x^7, x^5, x^3; cos: x^8, x^6, x^4), with the same signs
    for v, k in zip(c[4:7], (7, 5, 3)):
        assert (-1) ** ((k - 1) // 2) * v > 0 and abs(abs(v) * math.factorial(k) - 1) < 0.02
    for v, k in zip(c[7:10], (8, 6, 4)):
        assert (-1) ** (k // 2) * v > 0 and abs(abs(v) * math.factorial(k) - 1) < 0.02
    assert c[10] == -0.5 and c[11] == 1.0


def test_polynomials_within_1_5_ulp_on_the_reduced_interval():
    r = np.linspace(-math.pi / 4, math.pi / 4, 200001).astype(F)
    worst_s = worst_c = 0.0
    for x in r[::7]:
        sn, cs = _poly(x)
        es, ec = math.sin(float(x)), math.cos(float(x))
        if es != 0.0:
            worst_s = max(worst_s, abs(float(sn) - es) / float(np.spacing(F(abs(es)))))
        worst_c = max(worst_c, abs(float(cs) - ec) / float(np.spacing(F(abs(ec)))))
    assert worst_s < 1.5 and worst_c < 1.5, (worst_s, worst_c)


@pytest.mark.parametrize("lo,hi", [(-8192, 8192), (-10, 10), (0, 0.1)])
def test_fast_sincos_accuracy(lo, hi):
    xs = np.random.default_rng(int(hi)).uniform(lo, hi, 4000).astype(F)
    xs = np.concatenate([xs, (np.arange(-16, 17) * np.pi / 4).astype(F), np.array([0.0, -0.0], F)])
    err = 0.0
    for x in xs:
        s, c = fast_sincos(x)
        err = max(err, abs(float(s) - math.sin(float(x))), abs(float(c) - math.cos(float(x))))
    assert err < 1.2e-7   # about one float32 ulp of 1
