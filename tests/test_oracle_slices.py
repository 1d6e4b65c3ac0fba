"""Pins of the potential-slice oracle (CPU): explicit DFT matrices, a single
atom's closed form, linearity, and float32 vs float64 agreement."""
import numpy as np
import pytest

from oracle import emsim_oracle as orc

ZS = [1, 6, 8, 79]


def dft_mat(n, sign=-1):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def brute_slices(hist, sf_half, bandlimit):
    """V = Re IDFT2[ D sum_e F_e DFT2(h_e) ] with explicit matrices (float64)."""
    n_e, n_s, n1, n2 = hist.shape
    a, b = dft_mat(n1), dft_mat(n2)
    ai, bi = dft_mat(n1, +1), dft_mat(n2, +1)
    k1 = ((np.arange(n1) + n1 // 2) % n1) - n1 // 2
    k2 = ((np.arange(n2) + n2 // 2) % n2) - n2 // 2
    # full-plane form factor from the half table: F(k1, -k2) = F(-k1, k2)
    neg = np.arange(n2) > n2 // 2
    sf_full = np.empty((n_e, n1, n2))
    for e in range(n_e):
        sf_full[e][:, ~neg] = sf_half[e]
        sf_full[e][:, neg] = sf_half[e][(-np.arange(n1)) % n1][:, (n2 - np.arange(n2))[neg]]
    d = (k1[:, None] ** 2 + k2[None, :] ** 2 <= (min(n1, n2) // 2) ** 2) if bandlimit else 1
    out = np.empty((n_s, n1, n2))
    for s in range(n_s):
        acc = sum(a @ hist[e, s] @ b.T * sf_full[e] for e in range(n_e))
        out[s] = (ai @ (acc * d) @ bi.T).real / (n1 * n2)
    return np.clip(out, 1e-13, None)


@pytest.mark.parametrize("n,bandlimit", [(8, True), (16, True), (16, False)])
def test_slices_against_explicit_dft(n, bandlimit):
    rng = np.random.default_rng(n)
    hist = rng.integers(0, 3, size=(len(ZS), 3, n, n)).astype(np.float32)
    sf = orc.scattering_factors_half(ZS, n, n, 0.7, 0.7)
    ref = brute_slices(hist.astype(np.float64), sf.astype(np.float64), bandlimit)
    v64 = orc.make_multi_slices_f64(hist, sf, bandlimit)
    v32 = orc.make_multi_slices_fftw32(hist, sf, bandlimit)
    assert orc.rel_l2(v64, ref) < 1e-12
    assert orc.rel_l2(v32, ref) < 2e-6


def test_single_atom_at_origin_is_inverse_transform_of_form_factor():
    n = 32
    sf = orc.scattering_factors_half([6], n, n, 0.5, 0.5)
    hist = np.zeros((1, 1, n, n), np.float32)
    hist[0, 0, 0, 0] = 1.0
    v = orc.make_multi_slices_f64(hist, sf, symmetric_bandlimit=False)[0]
    expect = np.fft.irfft2(sf[0].astype(np.float64), s=(n, n))
    np.testing.assert_allclose(v, np.clip(expect, 1e-13, None), rtol=1e-12, atol=1e-15)
    # translating the atom translates the slice (periodic convolution)
    hist2 = np.roll(hist, (3, 5), axis=(2, 3))
    v2 = orc.make_multi_slices_f64(hist2, sf, symmetric_bandlimit=False)[0]
    np.testing.assert_allclose(v2, np.roll(v, (3, 5), axis=(0, 1)), rtol=1e-12, atol=1e-15)


def test_linearity_in_histogram_before_clip():
    n = 16
    rng = np.random.default_rng(1)
    sf = orc.scattering_factors_half(ZS, n, n, 1.0, 1.0)
    h1 = rng.integers(0, 2, size=(4, 2, n, n)).astype(np.float32)
    h2 = rng.integers(0, 2, size=(4, 2, n, n)).astype(np.float32)
    # all-positive potentials stay far above the clip, so V is linear in h
    v1 = orc.make_multi_slices_f64(h1 + 1, sf)
    v2 = orc.make_multi_slices_f64(h2 + 1, sf)
    v12 = orc.make_multi_slices_f64(h1 + h2 + 2, sf)
    assert orc.rel_l2(v12, v1 + v2) < 1e-12


def test_bandlimit_mask_is_disk_of_radius_n_over_2():
    m = orc.bandlimit_mask_half(16, 16)
    assert m.shape == (16, 9)
    k1 = ((np.arange(16) + 8) % 16) - 8
    k2 = np.arange(9)
    np.testing.assert_array_equal(m, (k1[:, None] ** 2 + k2[None, :] ** 2 <= 64).astype(m.dtype))
    with pytest.raises(ValueError):
        orc.bandlimit_mask_half(16, 8)


def test_fftw32_and_f64_agree_and_dtype_chain():
    n = 64
    rng = np.random.default_rng(2)
    sf = orc.scattering_factors_half(ZS, n, n, 1.0, 1.0)
    hist = (rng.random((4, 5, n, n)) < 0.01).astype(np.float32)
    v32 = orc.make_multi_slices_fftw32(hist, sf)
    v64 = orc.make_multi_slices_f64(hist, sf)
    assert v32.dtype == np.float64          # numpy 2.x band-limit promotes, as in the reference
    assert orc.rel_l2(v32, v64) < 2e-6
    v32_nb = orc.make_multi_slices_fftw32(hist, sf, symmetric_bandlimit=False)
    assert v32_nb.dtype == np.float32
    assert v32.min() >= 1e-13


def test_one_slice_equals_multi_slice_with_one_plane():
    n = 32
    rng = np.random.default_rng(5)
    sf = orc.scattering_factors_half(ZS, n, n, 0.4, 0.4)
    hist = (rng.random((4, n, n)) < 0.05).astype(np.float32)
    a = orc.make_one_slice_fftw32(hist, sf)
    b = orc.make_multi_slices_fftw32(hist[:, None], sf)[0]
    assert orc.rel_l2(a, b) < 2e-6


@pytest.mark.parametrize("n,bandlimit", [(8, True), (16, False)])
def test_numpy_backend_slices_against_explicit_dft(n, bandlimit):
    """The numpy backend's multi-slice build (slice_builder_numpy.py:39-45):
    same potential as the explicit-DFT brute force (float32 spectra, so
    1e-5), and for one slice exactly the one-slice builder's operations."""
    rng = np.random.default_rng(100 + n)
    hist = rng.integers(0, 3, size=(len(ZS), 4, n, n)).astype(np.float32)
    sf = orc.scattering_factors_half(ZS, n, n, 0.5, 0.5)
    got = orc.make_multi_slices_numpy(hist, sf, bandlimit)
    ref = brute_slices(hist.astype(np.float64), sf.astype(np.float64), bandlimit)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
    one = orc.make_one_slice_fftw32(hist[:, 1], sf, bandlimit)
    assert np.array_equal(orc.make_multi_slices_numpy(hist[:, 1:2], sf, bandlimit)[0], one)
