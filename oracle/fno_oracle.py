"""Fourier neural operator of the learned coarse-slice correction, restated in
numpy -- TEST INFRASTRUCTURE ONLY (the HIP kernels are the product; this is
the checker).

The network has no counterpart in the reference.  Its definition follows the
FNO of Li et al. (2021, "Fourier Neural Operator for Parametric PDEs"):
lifting P, L layers v <- act(W v + b + K v) with K v = irfft2(R . rfft2(v)
restricted to the retained modes), no activation after the last layer, and a
linear projection Q to two channels (real / imaginary part of the
correction).  Retained modes: k1 in [0, m) U [n - m, n), k2 in [0, m).

Two evaluations of the same network:
  * ``forward(..., emulate_bf16=False)``: float64 throughout.  Its spectral
    layer is written as products with cos/sin bases (the way the GPU computes
    it) and is pinned against ``spectral_conv_fft`` (numpy's rfft2/irfft2) in
    tests/test_fno_oracle.py; ``backward`` is its exact adjoint, pinned by
    finite differences.
  * ``forward(..., emulate_bf16=True)``: rounds every GEMM operand to bf16
    (nearest even) and every stored float32 intermediate to float32 at the
    points where the HIP kernels do, for a tight forward comparison.

Parameters are a dict: lift_w (cin_pad, C), lift_b (C,), w[l] (C, C),
b[l] (C,), r[l] (2m, m, C, C) complex, proj_w (C, 2), proj_b (2,).
Activations are channels-last (B, n, n, C).
"""
import numpy as np
from scipy.special import erf


def bf16_round(x) -> np.ndarray:
    """float64 -> float32 -> bf16 (round to nearest even) -> float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, u, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def f32_round(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def k1_values(n: int, modes: int) -> np.ndarray:
    j = np.arange(2 * modes)
    return np.where(j < modes, j, n - 2 * modes + j)


def bases(n: int, modes: int, bf16: bool = False):
    """(Fy, Fx2, Gx, Gy) of the truncated forward / inverse transforms.

    Fy (2m, n):   rows k2 = cos(2 pi k2 y / n), rows m + k2 = -sin(...)
    Fx2 (4m, n):  rows j = cos(2 pi k1_j x / n), rows 2m + j = sin(...)
    Gx (2n, 4m):  row 2x = [cos | -sin], row 2x + 1 = [sin | cos]  (inverse over k1)
    Gy (n, 2m):   [c cos | -c sin] / n^2, c = 1 for k2 = 0 else 2 (inverse rfft)
    """
    M1, M2 = 2 * modes, modes
    y = np.arange(n)
    k2 = np.arange(M2)
    ph = 2 * np.pi * ((k2[:, None] * y[None, :]) % n) / n          # (m, n)
    Fy = np.concatenate([np.cos(ph), -np.sin(ph)], axis=0)
    c = np.where(k2 == 0, 1.0, 2.0) / float(n * n)
    Gy = np.concatenate([(c[:, None] * np.cos(ph)).T, (-c[:, None] * np.sin(ph)).T], axis=1)
    kv = k1_values(n, modes)
    th = 2 * np.pi * ((kv[:, None] * y[None, :]) % n) / n          # (2m, n)
    Fx2 = np.concatenate([np.cos(th), np.sin(th)], axis=0)
    Gx = np.zeros((2 * n, 2 * M1))
    Gx[0::2, :M1] = np.cos(th).T
    Gx[0::2, M1:] = -np.sin(th).T
    Gx[1::2, :M1] = np.sin(th).T
    Gx[1::2, M1:] = np.cos(th).T
    if bf16:
        Fy, Fx2, Gx, Gy = (bf16_round(a) for a in (Fy, Fx2, Gx, Gy))
    return Fy, Fx2, Gx, Gy


def spectral_conv_fft(v: np.ndarray, r: np.ndarray, modes: int) -> np.ndarray:
    """K v through numpy's FFT: irfft2(R . rfft2(v) on the retained modes)."""
    B, n, _, C = v.shape
    V = np.fft.rfft2(v, axes=(1, 2))
    kv = k1_values(n, modes)
    S = V[:, kv, :modes, :]                                        # (B, 2m, m, C)
    O = np.einsum("bkmi,kmio->bkmo", S, r)
    Z = np.zeros((B, n, n // 2 + 1, C), dtype=np.complex128)
    Z[:, kv, :modes, :] = O
    return np.fft.irfft2(Z, s=(n, n), axes=(1, 2))


def _mix(PQ, r, M1, M2, C):
    """PQ (B, 2 M1, 2 M2 C) -> (S (B, M1, M2, C) complex, Ob (B, 2 M1, M2 C))."""
    B = PQ.shape[0]
    P = PQ[:, :M1].reshape(B, M1, 2, M2, C)
    Q = PQ[:, M1:].reshape(B, M1, 2, M2, C)
    S = (P[:, :, 0] + Q[:, :, 1]) + 1j * (P[:, :, 1] - Q[:, :, 0])
    O = np.einsum("bkmi,kmio->bkmo", S, r)
    Ob = np.concatenate([O.real.reshape(B, M1, M2 * C), O.imag.reshape(B, M1, M2 * C)], axis=1)
    return S, Ob


def forward(p: dict, x: np.ndarray, modes: int, emulate_bf16: bool = False):
    """x (B, n, n, cin_pad) -> out (B, n, n, 2), cache for ``backward``."""
    rb = bf16_round if emulate_bf16 else (lambda a: a)
    rf = f32_round if emulate_bf16 else (lambda a: a)
    x = np.asarray(x, dtype=np.float64)
    B, n, _, _ = x.shape
    C = p["lift_w"].shape[1]
    L = len(p["w"])
    M1, M2 = 2 * modes, modes
    Fy, Fx2, Gx, Gy = bases(n, modes, bf16=emulate_bf16)
    v = rb(rb(x) @ rb(p["lift_w"]) + p["lift_b"])
    cache = {"x": x, "v": [v], "z": [], "pq": [], "S": []}
    for l in range(L):
        T = rf(np.einsum("jy,bxyc->bxjc", Fy, v))                  # (B, n, 2 M2, C)
        PQ = rf(np.einsum("jx,bxm->bjm", Fx2, rb(T.reshape(B, n, 2 * M2 * C))))
        S, Ob = _mix(PQ, p["r"][l], M1, M2, C)
        Ob = rf(Ob)
        U = rf(np.einsum("xj,bjm->bxm", Gx, rb(Ob)))                # (B, 2n, M2 C)
        Uhat = U.reshape(B, n, 2 * M2, C)                          # rows (ri M2 + k2)
        z = rb(v) @ rb(p["w"][l]) + np.einsum("yj,bxjc->bxyc", Gy, rb(Uhat)) + p["b"][l]
        act = l + 1 < L
        v = rb(gelu(z) if act else z)
        cache["z"].append(rb(z))
        cache["pq"].append(PQ)
        cache["S"].append(S)
        cache["v"].append(v)
    out = rb(v) @ rb(p["proj_w"]) + p["proj_b"]
    return out, cache


def backward(p: dict, cache: dict, dout: np.ndarray, modes: int) -> dict:
    """Exact adjoint of the float64 forward: gradients of sum(out * G) for dout = G."""
    x = cache["x"]
    B, n, _, _ = x.shape
    C = p["lift_w"].shape[1]
    L = len(p["w"])
    M1, M2 = 2 * modes, modes
    Fy, Fx2, Gx, Gy = bases(n, modes)
    g = {"w": [None] * L, "b": [None] * L, "r": [None] * L}
    vL = cache["v"][L]
    g["proj_w"] = np.einsum("bxyc,bxyo->co", vL, dout)
    g["proj_b"] = dout.sum(axis=(0, 1, 2))
    dv = dout @ p["proj_w"].T
    for l in reversed(range(L)):
        z = cache["z"][l]
        dz = dv * gelu_grad(z) if l + 1 < L else dv
        vin = cache["v"][l]
        g["w"][l] = np.einsum("bxyi,bxyo->io", vin, dz)
        g["b"][l] = dz.sum(axis=(0, 1, 2))
        dUhat = np.einsum("yj,bxyc->bxjc", Gy, dz)                  # (B, n, 2 M2, C)
        dU = dUhat.reshape(B, 2 * n, M2 * C)
        dOb = np.einsum("xj,bxm->bjm", Gx, dU)                      # (B, 2 M1, M2 C)
        dO = (dOb[:, :M1] + 1j * dOb[:, M1:]).reshape(B, M1, M2, C)
        S = cache["S"][l]
        r = p["r"][l]
        # O = R S (complex-linear): dS = sum_o conj(R) dO, dR = sum_b conj(S) dO
        # (real-pair gradients of the real loss, written as complex numbers)
        dS = np.einsum("kmio,bkmo->bkmi", np.conj(r), dO)
        g["r"][l] = np.einsum("bkmi,bkmo->kmio", np.conj(S), dO)
        dP = np.zeros((B, M1, 2, M2, C))
        dQ = np.zeros((B, M1, 2, M2, C))
        dP[:, :, 0] = dS.real
        dQ[:, :, 1] = dS.real
        dP[:, :, 1] = dS.imag
        dQ[:, :, 0] = -dS.imag
        dPQ = np.concatenate([dP.reshape(B, M1, -1), dQ.reshape(B, M1, -1)], axis=1)
        dT = np.einsum("jx,bjm->bxm", Fx2, dPQ).reshape(B, n, 2 * M2, C)
        dv = dz @ p["w"][l].T + np.einsum("jy,bxjc->bxyc", Fy, dT)
    g["lift_w"] = np.einsum("bxyi,bxyc->ic", x, dv)
    g["lift_b"] = dv.sum(axis=(0, 1, 2))
    return g


def random_params(cin: int, channels: int, modes: int, layers: int, seed: int,
                  proj_scale: float = 1.0) -> dict:
    """Random parameters of the right shapes for the tests (uniform dense
    weights of scale 1/sqrt(fan_in), spectral weights (1/C^2) U(0, 1))."""
    rng = np.random.default_rng(seed)
    cin_pad = (cin + 15) // 16 * 16
    C = channels

    def uni(shape, fan_in):
        s = 1.0 / np.sqrt(fan_in)
        return rng.uniform(-s, s, size=shape)

    lift_w = np.zeros((cin_pad, C))
    lift_w[:cin] = uni((cin, C), cin)
    p = {"lift_w": lift_w, "lift_b": uni((C,), cin), "w": [], "b": [], "r": []}
    scale = 1.0 / (C * C)
    for _ in range(layers):
        p["w"].append(uni((C, C), C))
        p["b"].append(uni((C,), C))
        p["r"].append(scale * (rng.uniform(0, 1, (2 * modes, modes, C, C)) +
                               1j * rng.uniform(0, 1, (2 * modes, modes, C, C))))
    p["proj_w"] = proj_scale * uni((C, 2), C)
    p["proj_b"] = np.zeros(2)
    return p
