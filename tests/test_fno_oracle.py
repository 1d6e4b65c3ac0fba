"""Pins of the FNO oracle (oracle/fno_oracle.py), CPU only.

The network has no counterpart in the reference, so what pins it is
mathematics: the basis form of the spectral layer against numpy's FFT
definition, the backward pass against central finite differences of the
forward pass, and the bf16 rounding against torch's bfloat16 conversion.
"""
import copy

import numpy as np
import pytest
import torch

from oracle import fno_oracle as fo

N, C, M, L, B, CIN = 32, 16, 4, 3, 2, 5


@pytest.fixture(scope="module")
def setup():
    p = fo.random_params(CIN, C, M, L, seed=1)
    rng = np.random.default_rng(0)
    x = np.zeros((B, N, N, 16))
    x[..., :CIN] = rng.normal(size=(B, N, N, CIN))
    return p, x, rng


def test_spectral_layer_basis_form_equals_fft(setup):
    p, _, rng = setup
    v = rng.normal(size=(B, N, N, C))
    Fy, Fx2, Gx, Gy = fo.bases(N, M)
    T = np.einsum("jy,bxyc->bxjc", Fy, v)
    PQ = np.einsum("jx,bxm->bjm", Fx2, T.reshape(B, N, -1))
    _, Ob = fo._mix(PQ, p["r"][0], 2 * M, M, C)
    U = np.einsum("xj,bjm->bxm", Gx, Ob)
    K = np.einsum("yj,bxjc->bxyc", Gy, U.reshape(B, N, 2 * M, C))
    ref = fo.spectral_conv_fft(v, p["r"][0], M)
    assert np.abs(K - ref).max() < 1e-12 * np.abs(ref).max() + 1e-15


def test_spectral_layer_is_band_limited(setup):
    """A plane wave outside the retained modes is annihilated; one inside is
    multiplied by R at its mode."""
    p, _, _ = setup
    x = np.arange(N)
    hi = np.cos(2 * np.pi * (M + 2) * x / N)[None, :, None, None] * np.ones((1, N, N, C))
    assert np.abs(fo.spectral_conv_fft(hi, p["r"][0], M)).max() < 1e-12
    lo = np.ones((1, N, N, C))                        # k = 0 mode only
    out = fo.spectral_conv_fft(lo, p["r"][0], M)
    expect = np.real(np.ones(C) @ p["r"][0][0, 0])    # sum_i R[0,0,i,o]
    assert np.allclose(out[0, 3, 5], expect, rtol=1e-12)


def _loss(p, x, G):
    return float((fo.forward(p, x, M)[0] * G).sum())


@pytest.mark.parametrize("key", ["lift_w", "lift_b", "proj_w", "proj_b", "w", "b", "r"])
def test_backward_matches_finite_differences(setup, key):
    p, x, rng = setup
    out, cache = fo.forward(p, x, M)
    G = rng.normal(size=out.shape)
    g = fo.backward(p, cache, G, M)
    layers = range(L) if key in ("w", "b", "r") else [None]
    eps = 1e-6
    for l in layers:
        base = p[key][l] if l is not None else p[key]
        d = rng.normal(size=base.shape)
        if np.iscomplexobj(base):
            d = d + 1j * rng.normal(size=base.shape)
        pp, pm = copy.deepcopy(p), copy.deepcopy(p)
        if l is None:
            pp[key], pm[key] = base + eps * d, base - eps * d
            gk = g[key]
        else:
            pp[key][l], pm[key][l] = base + eps * d, base - eps * d
            gk = g[key][l]
        fd = (_loss(pp, x, G) - _loss(pm, x, G)) / (2 * eps)
        an = float(np.sum(gk.real * d.real) + np.sum(gk.imag * d.imag))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (key, l, fd, an)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=4096) * 10.0 ** rng.integers(-6, 6, 4096),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 65504.0]])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(fo.bf16_round(x).astype(np.float32), ref)


def test_emulated_forward_is_close_to_float64(setup):
    p, x, _ = setup
    a, _ = fo.forward(p, x, M)
    b, _ = fo.forward(p, x, M, emulate_bf16=True)
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-5 < rel < 3e-2
