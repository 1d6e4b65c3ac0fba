"""Pins of the host physics and the form-factor table (CPU)."""
import math

import numpy as np
import pytest

from emsim_amd import elem, physics
from oracle import emsim_oracle as orc


def test_interaction_parameter_identity():
    # restates the reference's only physics assertion (test/physics_test.py:8-16)
    e = 200
    lam = physics.electron_wave_length_angstrom(e)
    gamma = physics.electron_relativity_gamma(e)
    sigma = physics.interaction_parameter(e)
    assert abs(lam * gamma - sigma * 2 * math.pi * physics.a0 * physics.e * 1e8) < 1e-5


def test_wavelength_and_gamma_200kev():
    lam = physics.electron_wave_length_angstrom(200)
    gamma = physics.electron_relativity_gamma(200)
    # textbook values (Kirkland 2010, Table 2.2: 200 keV -> 0.02508 A, gamma 1.3914)
    assert abs(lam - 0.0250793) < 2e-7
    assert abs(gamma - 1.391390) < 2e-6
    lo, go = orc.wavelength_and_gamma(200)
    assert lo == lam and go == gamma


def test_scherzer_closed_forms():
    lam = physics.electron_wave_length_angstrom(200)
    df, aper = physics.scherzer_condition(lam, 1.3)
    assert abs(df - math.sqrt(1.5 * 1.3e7 * lam)) < 1e-9
    assert abs(aper - (6 * lam / 1.3e7) ** 0.25) < 1e-15
    assert abs(df - 699.319) < 1e-2 and abs(aper - 10.3724e-3) < 1e-6


def test_kirkland_table_matches_reference_layout():
    p = elem.elem_params([1, 6, 92])
    assert p.shape == (3, 13) and p.dtype == np.float32
    # Z=1 row of the reference data file, first line of parameters:
    # a1 b1 a2 b2 = 4.20298324e-003 2.25350888e-001 6.27762505e-002 2.25366950e-001
    assert p[0, 1] == np.float32(4.20298324e-003) and p[0, 4] == np.float32(2.25350888e-001)
    assert p[0, 2] == np.float32(6.27762505e-002) and p[0, 5] == np.float32(2.25366950e-001)
    with pytest.raises(ValueError):
        elem.elem_params([0])


def test_element_symbols():
    assert [elem.number(s) for s in ["H", "He", "Li", "Be", "B"]] == [1, 2, 3, 4, 5]
    assert elem.number("au") == 79 and elem.number("Xx") is None


def test_form_factor_at_zero_is_closed_form():
    zs = [1, 6, 14, 79]
    f = elem.scattering_factor_function(zs)(np.zeros(1))
    p = elem.elem_params(zs).astype(np.float64)
    expect = (p[:, 1:4] / p[:, 4:7]).sum(1) + p[:, 7:10].sum(1)
    np.testing.assert_allclose(f[:, 0], expect, rtol=1e-6)


def test_form_factor_pointwise_against_scalar_formula():
    zs = [8, 29]
    q = np.array([0.0, 0.1, 0.5, 1.7, 3.99, 4.0, 5.0])
    f = elem.scattering_factor_function(zs)(q)
    p = elem.elem_params(zs)
    for k in range(2):
        for i, qq in enumerate(q):
            q2 = qq * qq
            s = sum(float(p[k, 1 + m]) / (q2 + float(p[k, 4 + m]))
                    + float(p[k, 7 + m]) * math.exp(-float(p[k, 10 + m]) * q2) for m in range(3))
            expect = s if q2 < 16 else 0.0
            assert f[k, i] == np.float32(expect) or abs(f[k, i] - expect) <= 1e-6 * abs(expect)


@pytest.mark.parametrize("n1,n2,d1,d2", [(16, 16, 1.0, 1.0), (32, 16, 0.5, 0.25),
                                         (512, 512, 50 / 512, 50 / 512)])
def test_half_table_oracle_equals_package(n1, n2, d1, d2):
    zs = [1, 6, 7, 8, 16]
    a = elem.scattering_factors_half(zs, n1, n2, d1, d2)
    b = orc.scattering_factors_half(zs, n1, n2, d1, d2)
    assert a.shape == (5, n1, n2 // 2 + 1) and a.dtype == np.float32
    np.testing.assert_array_equal(a, b)
    # FFT order: entry [.., 0, 0] is q = 0, the largest value
    assert np.all(a[:, 0, 0] == a.reshape(5, -1).max(1))
    # symmetric in +-k1 (real, even form factors)
    np.testing.assert_array_equal(a[:, 1:, :], a[:, :0:-1, :])


def test_form_factor_is_the_hankel_transform_of_the_projected_potential():
    """The oracle's f_e(q) table against the real-space projected potential
    of the same Kirkland fit (emsim/elem.py:67-117, Kirkland eq. 5.10):
        v_z(r) = 2 pi a0 e [2 pi sum a_i K0(2 pi r sqrt b_i)
                            + pi sum c_i / d_i exp(-pi^2 r^2 / d_i)],
    whose 2-D Fourier transform, 2 pi int v_z(r) J0(2 pi q r) r dr, is
    2 pi a0 e f_e(q).  The transform is integrated numerically, so a swapped
    or misread parameter column, a dropped term or a wrong q scale in the
    table fails here."""
    special = pytest.importorskip("scipy.special")
    integrate = pytest.importorskip("scipy.integrate")
    zs = [1, 6, 8, 16, 79]
    n, d = 64, 0.25                            # table frequencies j / (n d) = j / 16
    sf = orc.scattering_factors_half(zs, n, n, d, d)
    p = orc.kirkland_params(zs).astype(np.float64)
    edges = np.linspace(0.0, 60.0, 241)
    for k in range(len(zs)):
        a, b, c, dd = p[k, 1:4], p[k, 4:7], p[k, 7:10], p[k, 10:13]

        def v(r):
            return (2 * math.pi * sum(a[i] * special.k0(2 * math.pi * r * math.sqrt(b[i]))
                                      for i in range(3))
                    + math.pi * sum(c[i] / dd[i] * math.exp(-math.pi ** 2 * r * r / dd[i])
                                    for i in range(3)))
        for j in (0, 2, 8, 16, 32):
            q = j / (n * d)
            f = lambda r: 2 * math.pi * v(r) * special.j0(2 * math.pi * q * r) * r
            val = sum(integrate.quad(f, lo, hi, limit=200)[0]
                      for lo, hi in zip(edges[:-1], edges[1:]))
            assert abs(val - float(sf[k, 0, j])) <= 1e-6 * val, (zs[k], q)
            assert sf[k, j, 0] == sf[k, 0, j]      # radial: the same q along either axis


def test_form_factor_lorentzian_weights_follow_the_mott_asymptote():
    """At large q the Mott-Bethe formula gives f_e(q) -> Z / (2 pi^2 a0 q^2),
    so the Lorentzian weights of the fit sum to about Z / (2 pi^2 a0) (a0 in
    Angstrom); Kirkland's fits meet it to a few per cent."""
    zs = [1, 6, 7, 8, 16, 29, 79]
    p = orc.kirkland_params(zs).astype(np.float64)
    a0_angstrom = physics.a0 * 1e8
    for z, row in zip(zs, p):
        ratio = row[1:4].sum() / (z / (2 * math.pi ** 2 * a0_angstrom))
        assert abs(ratio - 1) < 0.04, (z, ratio)
