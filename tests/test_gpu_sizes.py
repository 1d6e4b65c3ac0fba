"""GPU parity at lateral sizes that are not powers of two.

The reference's FFTW plans take any size (emsim/backend/fftw_ext/src/
WavePropagator.cpp:25-33, SliceBuilder.cpp:41-55), and pot.build_multi_slices
without lateral_size sizes the grid from the molecule's extent
(emsim/pot.py:51-74, atoms.py:214-223).  The hip backend runs those sizes
through chirp-z transforms on its power-of-two FFT kernels; each call is
checked against the oracle at even and odd sizes.
"""
import numpy as np
import pytest
import torch

import emsim_amd
from emsim_amd import atoms as atm
from emsim_amd import pot
from emsim_amd import synth
from oracle import emsim_oracle as orc
from tests._chain import slices_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5
LAM, GAMMA = orc.wavelength_and_gamma(200.0)


@pytest.fixture(autouse=True)
def _gpu(hip):
    emsim_amd.config.set_backend("hip")
    yield
    torch.cuda.synchronize()


def _host(x):
    return x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _wave(n1, n2, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n1, n2)) + 1j * rng.standard_normal((n1, n2))).astype(
        np.complex64)


def test_size_support_flags(hip):
    assert hip.fft_size_supported(512) and hip.fft_size_supported(90)
    assert hip.raw("emsim_hip_fft_size_supported")(90) == 2
    assert not hip.fft_size_supported(4) and not hip.fft_size_supported(3000)


@pytest.mark.parametrize("n1,n2,d", [(90, 90, 1.0), (125, 125, 0.5), (96, 200, 0.8),
                                     (300, 256, 1.0)])
def test_space_and_lens_propagate(n1, n2, d):
    prop = emsim_amd.wave.get_raw_wave_propagator((n1, n2), d, 200)
    w = _wave(n1, n2, n1 + n2)
    got = _host(prop.space_propagate(w, 2.0))
    ref = orc.space_propagate_fftw32(w, 2.0, LAM, d, d)
    e = orc.rel_l2(got, ref)
    got_l = _host(prop.lens_propagate(w, 1.3, 700.0, 10e-3))
    ref_l = orc.lens_propagate_fftw32(w, LAM, 1.3, 700.0, 10e-3, d, d)
    el = orc.rel_l2(got_l, ref_l)
    print(f"[{n1}x{n2}] space_propagate {e:.2e}, lens_propagate {el:.2e}", flush=True)
    assert e < TOL and el < TOL


@pytest.mark.parametrize("n,ns", [(75, 6), (150, 10)])
def test_multislice_and_slices(n, ns):
    """Potential slices (band-limited, square grid) and multislice of a
    protein-like cloud at an odd and an even size."""
    mol = atm.centralize(synth.protein_cloud(3000, 400 + n, diameter=0.7 * n))
    got_v = _host(pot.build_multi_slices(mol, dz=2.0, pixel_size=1.0, lateral_size=n,
                                         n_slices=ns))
    v32, v64, _ = slices_oracle(mol, 2.0, 1.0, n, ns)
    ev, ev64 = orc.rel_l2(got_v, v32), orc.rel_l2(got_v, v64)
    prop = emsim_amd.wave.get_raw_wave_propagator(n, 1.0, 200)
    w0 = orc.init_wave(n, n, 1.0, 1.0, 20.0)
    got_x = _host(prop.multislice_propagate(w0, v32.astype(np.float32), 2.0))
    ref_x = orc.multislice_fftw32(w0, v32.astype(np.float32), 2.0, LAM, GAMMA, 1.0, 1.0)
    ex = orc.rel_l2(got_x, ref_x)
    print(f"[{n} x {ns}] slices {ev:.2e} (fp64 {ev64:.2e}), multislice {ex:.2e}", flush=True)
    assert ev < TOL and ev64 < TOL and ex < TOL


def test_one_slice_odd_size():
    mol = atm.centralize(synth.five_atoms())
    pixel, n = 50 / 512, 511
    got = _host(pot.build_one_slice(mol, pixel_size=pixel, lateral_size=(n, n)))
    m, zs, counts = atm.sort_elements_and_count(mol)
    hist = orc.bin_atoms_one_slice(m.coordinates, counts, n, n, pixel, pixel)
    sf = orc.scattering_factors_half(list(zs), n, n, pixel, pixel)
    ref = orc.make_one_slice_fftw32(hist, sf)
    e = orc.rel_l2(got, ref)
    print(f"[one slice {n}] {e:.2e}", flush=True)
    assert e < TOL


def test_grid_sized_from_the_molecule():
    """pot.build_multi_slices without lateral_size sizes the grid from the
    molecule (ceil(extent / pixel) per axis, not a power of two); the slices,
    the propagator built for their shape, multislice, lens and |psi|^2 match
    the oracle chain.  (The reference's Pipe with roi=None fails with a
    TypeError before any of this, wave.py:10-16 indexing None; so does this
    Pipe.)"""
    tem = emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                            aperture=10e-3)
    cloud = synth.protein_cloud(6000, 410)
    # two H atoms at opposite corners square the x-y extent (49.5 A), so the
    # reference's band-limit, which needs n1 == n2, applies
    corners = np.array([[0.0, -24.75, -24.75], [0.0, 24.75, 24.75]], dtype=np.float32)
    mol = atm.centralize(atm.AtomList(np.concatenate([cloud.elements, [1, 1]]),
                                      np.concatenate([cloud.coordinates, corners])))
    v = pot.build_multi_slices(mol, dz=2.0, pixel_size=1.0)
    ns, n1, n2 = tuple(v.shape)
    assert n1 & (n1 - 1) and n2 & (n2 - 1)
    prop = emsim_amd.wave.get_raw_wave_propagator((n1, n2), 1.0, 200)
    exit_wave = prop.multislice_propagate(prop.init_wave(tem.electron_dose), v, 2.0)
    image = _host(prop.lens_intensity(exit_wave, tem.cs_mm, tem.defocus, tem.aperture))
    v32, v64, _ = slices_oracle(mol, 2.0, 1.0, None)
    assert v32.shape == (ns, n1, n2)
    w0 = orc.init_wave(n1, n2, 1.0, 1.0, tem.electron_dose)
    x32 = orc.multislice_fftw32(w0, v32.astype(np.float32), 2.0, LAM, GAMMA, 1.0, 1.0)
    i32 = orc.intensity(orc.lens_propagate_fftw32(x32, LAM, tem.cs_mm, tem.defocus,
                                                  tem.aperture, 1.0, 1.0))
    e = orc.rel_l2(image.reshape(i32.shape), i32)
    print(f"[grid from the molecule {n1}x{n2}, {ns} slices] image vs fftw32 {e:.2e}", flush=True)
    assert e < TOL
    pipe = emsim_amd.pipe.Pipe(tem, resolution=2.0, slice_thickness=2.0, roi=None,
                               add_water=False)
    with pytest.raises(TypeError):
        pipe.run(mol)


def test_batched_waves_equal_single_and_fused_path_refuses():
    prop = emsim_amd.wave.get_raw_wave_propagator(100, 1.0, 200)
    w = np.stack([_wave(100, 100, 1), _wave(100, 100, 2)])
    both = _host(prop.space_propagate(w, 3.0))
    for b in range(2):
        one = _host(prop.space_propagate(w[b], 3.0))
        np.testing.assert_array_equal(both[b].reshape(one.shape), one)
    from emsim_amd.batch import BatchPipe
    tem = emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                            aperture=10e-3)
    bp = BatchPipe(tem, 1.0, 2.0, 100, 4, [1, 6, 7, 8, 16])
    batch = bp.upload([bp.prepare(synth.protein_cloud(500, 3))])
    with pytest.raises(ValueError, match="power-of-two"):
        bp.run(batch)
