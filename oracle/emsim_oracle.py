"""CPU oracle for the emsim multislice path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker / CPU baseline.
The product (``emsim_amd``) never imports it.

Three restatements of the reference's algorithm live here, each function
citing the reference lines it follows (paths relative to the reference repo):

* ``*_fftw32``  -- the fftw backend's float32 semantics: binning via numpy
  float64 edges (``emsim/backend/slice_builder_base.py:180-225``), batched
  R2C -> sum over elements in float32 -> unnormalised C2R
  (``emsim/backend/fftw_ext/src/SliceBuilder.cpp:27-85``), the numpy
  band-limit + clip that follows it (``slice_builder_fftw.py:43-48``,
  ``slice_builder_base.py:166-175``), and the float32 wave kernels
  (``emsim/backend/fftw_ext/src/WavePropagator.cpp:41-138``).  FFTW itself is
  not available here; numpy's float32 pocketfft stands in for it (both are
  exact DFTs up to float32 rounding).
* ``*_f64`` -- the same algorithm in float64/complex128 (the FP64 oracle the
  float32 results are also measured against).  The discrete choices the
  reference makes in float32 (which frequencies the 2/3 mask and the
  aperture keep) are taken from the float32 computation so that both oracles
  describe the same algorithm.
* ``*_numpy`` -- the reference's numpy backend
  (``emsim/backend/wave_propagator_numpy.py``, ``slice_builder_numpy.py``),
  whose linspace frequency grid differs from the fftw backend's (reported,
  not hidden).

Parity pins (each checked in ``tests/test_oracle_*.py``): numpy.histogramdd
on the reference's own edges plus brute-force per-atom binning; explicit DFT
matrices at 8x8 / 16x16; closed forms (single atom at the origin, plane
wave, free-space composition, identity lens); norm conservation.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

F32 = np.float32
PI32 = np.float32(3.14159265358979324)

_KIRKLAND = Path(__file__).resolve().parent.parent / "emsim_amd" / "data" / "kirkland_params.txt"


# ----------------------------------------------------------------------------
# physics (emsim/physics.py:5-59,136-156)
# ----------------------------------------------------------------------------
def wavelength_and_gamma(beam_energy_kev: float):
    h = 6.62607015e-27
    c = 29979245800.0
    m0 = 9.1093835611e-28
    e_c = 1.602176634e-19
    m0c2_kev = m0 * c ** 2 * 1e-7 / e_c / 1000.
    hc_kev_ang = h * c * 0.01 / e_c
    lam = hc_kev_ang / math.sqrt(beam_energy_kev ** 2 + 2 * m0c2_kev * beam_energy_kev)
    gamma = 1 + beam_energy_kev / m0c2_kev
    return lam, gamma


# ----------------------------------------------------------------------------
# form factors (emsim/elem.py:120-156,267-291; slice_builder_base.py:30-33)
# ----------------------------------------------------------------------------
def kirkland_params(zs):
    raw = np.loadtxt(_KIRKLAND, dtype=np.float64, comments="#")
    return np.array([raw[int(z) - 1, 1:] for z in zs], dtype=F32)


def scattering_factors_half(zs, n1, n2, d1, d2):
    """(nE, n1, n2//2+1) float32 form-factor table in FFT order."""
    pms = kirkland_params(zs)
    kx = np.fft.fftshift(np.fft.fftfreq(n1, d1))
    ky = np.fft.fftshift(np.fft.fftfreq(n2, d2))
    kx, ky = np.meshgrid(kx, ky, indexing="ij")
    q = np.sqrt(kx * kx + ky * ky)
    q2 = q * q
    out = np.empty((len(zs), n1, n2), dtype=F32)
    for k in range(len(zs)):
        a, b, c, d = pms[k, 1:4], pms[k, 4:7], pms[k, 7:10], pms[k, 10:13]
        acc = np.sum([a[i] / (q2 + b[i]) + c[i] * np.exp(-d[i] * q2) for i in range(3)], axis=0)
        out[k] = np.where(q2 < 16, acc, 0.)
    out = np.fft.ifftshift(out, axes=(-2, -1))
    return np.ascontiguousarray(out[:, :, :n2 // 2 + 1], dtype=F32)


# ----------------------------------------------------------------------------
# binning (slice_builder_base.py:180-225)
# ----------------------------------------------------------------------------
def bin_edges(n: int, d: float) -> np.ndarray:
    """The reference's float64 edges: arange(s, s + n d + d, d), s = -d (n // 2)."""
    start = -d * (n // 2)
    end = start + n * d
    return np.arange(start, end + d, d)


def bin_atoms_multi_slices(coords, counts, n0, n1, n2, d0, d1, d2):
    """(nE, n0, n1, n2) float32 integer-valued histograms via numpy.histogramdd."""
    edges = [bin_edges(n, d) for n, d in zip((n0, n1, n2), (d0, d1, d2))]
    nb = [len(e) - 1 for e in edges]
    out = np.zeros((len(counts), *nb), dtype=F32)
    m = 0
    for i, n in enumerate(counts):
        n = int(n)
        if n:
            out[i], _ = np.histogramdd(coords[m:m + n], bins=edges)
        m += n
    return out


def bin_atoms_one_slice(coords, counts, n1, n2, d1, d2):
    """(nE, n1, n2) float32 histograms of columns 1, 2 via numpy.histogram2d."""
    edges = [bin_edges(n, d) for n, d in zip((n1, n2), (d1, d2))]
    nb = [len(e) - 1 for e in edges]
    out = np.zeros((len(counts), *nb), dtype=F32)
    x, y = coords[:, 1], coords[:, 2]
    m = 0
    for i, n in enumerate(counts):
        n = int(n)
        if n:
            out[i], _, _ = np.histogram2d(x[m:m + n], y[m:m + n], bins=edges)
        m += n
    return out


def bin_index_bruteforce(x: float, edges: np.ndarray) -> int:
    """Bin of one float32 coordinate by the searchsorted(side='right') rule.

    Returns -1 for dropped atoms.  The last edge is closed; NaN is dropped.
    """
    xv = float(x)
    if not (xv == xv):
        return -1
    nb = len(edges) - 1
    if xv == float(edges[-1]):
        return nb - 1
    cnt = 0
    for e in edges:
        if e <= xv:
            cnt += 1
    b = cnt - 1
    return b if 0 <= b < nb else -1


# ----------------------------------------------------------------------------
# symmetric band-limit (slice_builder_base.py:166-175)
# ----------------------------------------------------------------------------
def bandlimit_mask_half(n1: int, n2: int) -> np.ndarray:
    """The reference's disk mask restricted to the rfft half-plane.

    Only square grids broadcast in the reference; for odd n the mask is
    centred on arange(-n//2, -n//2 + n), i.e. shifted by one bin.
    """
    if n1 != n2:
        raise ValueError("the reference's symmetric band-limit only broadcasts for n1 == n2")
    r = min(n1, n2) // 2
    kx, ky = np.meshgrid(np.arange(-n1 // 2, -n1 // 2 + n1), np.arange(-n2 // 2, -n2 // 2 + n2))
    fil = np.where(kx ** 2 + ky ** 2 <= r ** 2, 1, 0)
    fil = np.fft.ifftshift(fil)
    return fil[:, :n2 // 2 + 1]


def symmetric_bandlimit_real(arr):
    """numpy 2.x dtype chain: float32 -> complex64 * int64 -> complex128 -> float64."""
    n1, n2 = arr.shape[-2:]
    fil = bandlimit_mask_half(n1, n2)
    ft = np.fft.rfft2(arr, axes=(-2, -1), s=(n1, n2))
    return np.fft.irfft2(ft * fil, s=(n1, n2))


# ----------------------------------------------------------------------------
# potential slices
# ----------------------------------------------------------------------------
def make_multi_slices_fftw32(hist, sf, symmetric_bandlimit=True):
    """fftw backend: SliceBuilder.cpp:27-85 then slice_builder_fftw.py:43-48."""
    n_e, n_s, n1, n2 = hist.shape
    npix = F32(n1 * n2)
    x = np.fft.rfft2(hist.astype(F32), axes=(-2, -1))          # complex64
    acc_re = np.zeros((n_s, n1, n2 // 2 + 1), dtype=F32)
    acc_im = np.zeros_like(acc_re)
    for k in range(n_e):                                         # element order, f32
        acc_re += x[k].real * sf[k][None]
        acc_im += x[k].imag * sf[k][None]
    y = (acc_re / npix) + 1j * (acc_im / npix)
    y = y.astype(np.complex64)
    v0 = np.fft.irfft2(y, s=(n1, n2), axes=(-2, -1), norm="forward").astype(F32)
    v = symmetric_bandlimit_real(v0) if symmetric_bandlimit else v0
    return np.clip(v, 1e-13, None)


def make_multi_slices_numpy(hist, sf, symmetric_bandlimit=True):
    """numpy backend (slice_builder_numpy.py:39-45): complex64 half spectra
    times the float32 table, numpy's sum over elements, normalised irfft2,
    band-limit (float64 under numpy 2.x), clip."""
    n_e, n_s, n1, n2 = hist.shape
    phase = np.fft.rfft2(hist.astype(F32), axes=(-2, -1))        # complex64
    phase *= sf[:, None, :, :]
    slices = np.fft.irfft2(np.sum(phase, axis=0), s=(n1, n2))
    if symmetric_bandlimit:
        slices = symmetric_bandlimit_real(slices)
    np.clip(slices, 1e-13, None, out=slices)
    return slices


def make_multi_slices_f64(hist, sf, symmetric_bandlimit=True):
    """FP64 oracle of the same potential build."""
    n_e, n_s, n1, n2 = hist.shape
    x = np.fft.rfft2(hist.astype(np.float64), axes=(-2, -1))
    y = np.einsum("esij,eij->sij", x, sf.astype(np.float64))
    if symmetric_bandlimit:
        y = y * bandlimit_mask_half(n1, n2)
    v = np.fft.irfft2(y, s=(n1, n2), axes=(-2, -1))
    return np.clip(v, 1e-13, None)


def make_one_slice_fftw32(hist, sf, symmetric_bandlimit=True):
    """One-slice builder of the fftw (== numpy) backend (slice_builder_fftw.py:24-32)."""
    n_e, n1, n2 = hist.shape
    phase = np.fft.rfft2(hist.astype(F32), axes=(-2, -1))        # complex64
    phase *= sf
    aslice = np.fft.irfft2(np.sum(phase, axis=0), s=(n1, n2))   # float32
    if symmetric_bandlimit:
        aslice = symmetric_bandlimit_real(aslice)
    np.clip(aslice, 1e-13, None, out=aslice)
    return aslice


def make_one_slice_f64(hist, sf, symmetric_bandlimit=True):
    return make_multi_slices_f64(hist[:, None], sf, symmetric_bandlimit)[0]


# ----------------------------------------------------------------------------
# wave propagation: fftw float32 semantics (WavePropagator.cpp)
# ----------------------------------------------------------------------------
def _freq_int(n: int) -> np.ndarray:
    """Signed integer frequency of FFT bin i, via the reference's ifftshift map."""
    i = np.arange(n)
    shifted = i + np.where(i < (n + 1) // 2, n // 2, -((n + 1) // 2))
    return shifted - n // 2


def freq2_f32(n1, n2, d1, d2):
    """f^2 on the FFT grid in float32, in WavePropagator.cpp:65-72's operation order."""
    dfx = F32(1.0) / F32(d1) / F32(n1)
    dfy = F32(1.0) / F32(d2) / F32(n2)
    fx = _freq_int(n1).astype(F32) * dfx
    fy = _freq_int(n2).astype(F32) * dfy
    return (fx * fx)[:, None] + (fy * fy)[None, :]


def propagator_mask_f32(n1, n2, d1, d2):
    d1f, d2f = F32(d1), F32(d2)
    fmax = F32(0.5) / (d1f if d1f >= d2f else d2f)
    filt = F32(0.6667) * fmax
    return freq2_f32(n1, n2, d1, d2) <= filt * filt


def init_wave(n1, n2, d1, d2, dose):
    """wave_propagator_base.py:25-29: constant sqrt(dose d1 d2), complex64."""
    w = np.ones((n1, n2), dtype=np.complex64)
    w *= np.sqrt(dose * d1 * d2) / np.abs(w)
    return w


def slice_transmit_fftw32(wave, aslice, lam, gamma):
    factor = F32(lam) * F32(gamma)
    ph = aslice.astype(F32) * factor
    t = np.cos(ph) + 1j * np.sin(ph)
    return (wave.astype(np.complex64) * t.astype(np.complex64)).astype(np.complex64)


def _fourier_multiply_fftw32(wave, mult, keep):
    n1, n2 = wave.shape
    w = np.fft.fft2(wave.astype(np.complex64))
    w = np.where(keep, (w * mult) / F32(n1 * n2), 0).astype(np.complex64)
    return np.fft.ifft2(w, norm="forward").astype(np.complex64)


def space_propagate_fftw32(wave, dz, lam, d1, d2):
    n1, n2 = wave.shape
    f2 = freq2_f32(n1, n2, d1, d2)
    arg = F32(lam) * PI32 * F32(dz) * f2
    p = (np.cos(arg) - 1j * np.sin(arg)).astype(np.complex64)
    return _fourier_multiply_fftw32(wave, p, propagator_mask_f32(n1, n2, d1, d2))


def multislice_fftw32(wave, slices, dz, lam, gamma, d1, d2):
    psi = wave
    for s in range(slices.shape[0]):
        psi = space_propagate_fftw32(slice_transmit_fftw32(psi, slices[s], lam, gamma), dz, lam, d1, d2)
    return psi


def lens_constants_f32(lam, cs_mm, defocus, aperture):
    lamf = F32(lam)
    c1 = F32(0.5) * PI32 * F32(cs_mm) * F32(1e7) * lamf * lamf * lamf
    c2 = PI32 * F32(defocus) * lamf
    f_aper = F32(aperture) / lamf
    return c1, c2, f_aper


def lens_propagate_fftw32(wave, lam, cs_mm, defocus, aperture, d1, d2):
    n1, n2 = wave.shape
    c1, c2, f_aper = lens_constants_f32(lam, cs_mm, defocus, aperture)
    f2 = freq2_f32(n1, n2, d1, d2)
    keep = f2 <= f_aper * f_aper
    chi = c1 * f2 * f2 - c2 * f2
    h = (np.cos(chi) - 1j * np.sin(chi)).astype(np.complex64)
    return _fourier_multiply_fftw32(wave, h, keep)


def intensity(wave):
    return (wave.real ** 2 + wave.imag ** 2).astype(F32)


# ----------------------------------------------------------------------------
# wave propagation: FP64 oracle of the same algorithm
# ----------------------------------------------------------------------------
def freq2_f64(n1, n2, d1, d2):
    fx = _freq_int(n1) / (d1 * n1)
    fy = _freq_int(n2) / (d2 * n2)
    return (fx * fx)[:, None] + (fy * fy)[None, :]


def slice_transmit_f64(wave, aslice, lam, gamma):
    return wave.astype(np.complex128) * np.exp(1j * (lam * gamma) * aslice.astype(np.float64))


def space_propagate_f64(wave, dz, lam, d1, d2):
    n1, n2 = wave.shape
    keep = propagator_mask_f32(n1, n2, d1, d2)
    p = np.exp(-1j * math.pi * lam * dz * freq2_f64(n1, n2, d1, d2))
    return np.fft.ifft2(np.where(keep, np.fft.fft2(wave.astype(np.complex128)) * p, 0))


def multislice_f64(wave, slices, dz, lam, gamma, d1, d2):
    psi = wave.astype(np.complex128)
    for s in range(slices.shape[0]):
        psi = space_propagate_f64(slice_transmit_f64(psi, slices[s], lam, gamma), dz, lam, d1, d2)
    return psi


def lens_propagate_f64(wave, lam, cs_mm, defocus, aperture, d1, d2):
    n1, n2 = wave.shape
    _, _, f_aper = lens_constants_f32(lam, cs_mm, defocus, aperture)
    keep = freq2_f32(n1, n2, d1, d2) <= f_aper * f_aper
    f2 = freq2_f64(n1, n2, d1, d2)
    chi = 0.5 * math.pi * cs_mm * 1e7 * lam ** 3 * f2 * f2 - math.pi * defocus * lam * f2
    return np.fft.ifft2(np.where(keep, np.fft.fft2(wave.astype(np.complex128)) * np.exp(-1j * chi), 0))


# ----------------------------------------------------------------------------
# numpy backend (wave_propagator_base.py:17-53, wave_propagator_numpy.py)
# ----------------------------------------------------------------------------
def numpy_backend_grid(n1, n2, d1, d2):
    qx = np.linspace(-0.5 / d1, 0.5 / d1, n1)
    qy = np.linspace(-0.5 / d2, 0.5 / d2, n2)
    qx, qy = np.meshgrid(qx, qy, indexing="ij")
    q = np.sqrt(qx * qx + qy * qy)
    fil = np.fft.ifftshift(np.where(q <= 0.5 / max(d1, d2) * 0.6667, 1., 0.))
    return q, fil


def slice_transmit_numpy(wave, aslice, lam, gamma):
    return wave * np.exp(1j * gamma * lam * aslice)


def multislice_numpy(wave, slices, dz, lam, gamma, d1, d2):
    n1, n2 = wave.shape
    q, fil = numpy_backend_grid(n1, n2, d1, d2)
    p = np.fft.ifftshift(np.exp(-1j * lam * np.pi * dz * q ** 2))
    t = np.exp(1j * gamma * lam * slices)
    psi = wave
    for s in range(slices.shape[0]):
        psi = np.fft.ifft2(p * np.fft.fft2(psi * t[s]) * fil)
    return psi


def lens_propagate_numpy(wave, lam, cs_mm, defocus, aperture, d1, d2):
    n1, n2 = wave.shape
    q, _ = numpy_backend_grid(n1, n2, d1, d2)
    cslambda = cs_mm * 1e7 * lam
    big_d = defocus / math.sqrt(cslambda)
    big_k = q * pow(cslambda * lam ** 2, 0.25)
    h = np.exp(-1j * (math.pi * (0.5 * big_k ** 4 - big_d * big_k ** 2)))
    h *= np.where(q < aperture / lam, 1., 0.)
    return np.fft.ifft2(np.fft.ifftshift(h) * np.fft.fft2(wave))


# ----------------------------------------------------------------------------
# relative L2 over a complex or real field
# ----------------------------------------------------------------------------
def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
