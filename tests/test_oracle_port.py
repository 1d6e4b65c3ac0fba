"""The C/OpenMP CPU baseline (oracle/fftw_port.c) against the numpy oracle."""
import numpy as np
import pytest

from emsim_amd import atoms as atm
from emsim_amd import synth
from oracle import emsim_oracle as orc


@pytest.mark.parametrize("n,ns,dz,pixel", [(64, 6, 2.0, 1.0), (128, 4, 1.0, 0.5)])
def test_port_matches_numpy_oracle(n, ns, dz, pixel):
    port = pytest.importorskip("oracle.port")
    try:
        port.load()
    except Exception as exc:     # no C compiler
        pytest.skip(str(exc))
    mol = atm.centralize(synth.protein_cloud(1500, 3, diameter=0.8 * n * pixel))
    mol, zs, counts = atm.sort_elements_and_count(mol)
    lam, gamma = orc.wavelength_and_gamma(200)
    sf = orc.scattering_factors_half(list(zs), n, n, pixel, pixel)
    hist = orc.bin_atoms_multi_slices(mol.coordinates, counts, ns, n, n, dz, pixel, pixel)
    v = orc.make_multi_slices_fftw32(hist, sf).astype(np.float32)
    w0 = orc.init_wave(n, n, pixel, pixel, 20.0)
    x = orc.multislice_fftw32(w0, v, dz, lam, gamma, pixel, pixel)
    ref = orc.intensity(orc.lens_propagate_fftw32(x, lam, 1.3, 700.0, 10e-3, pixel, pixel))
    got = port.simulate(mol.coordinates, counts, ns, n, dz, pixel, sf, lam, gamma, 1.3, 700.0,
                        10e-3, float(w0[0, 0].real))
    assert orc.rel_l2(got, ref) < 1e-5
    img, wave = port.simulate(mol.coordinates, counts, ns, n, dz, pixel, sf, lam, gamma, 1.3,
                              700.0, 10e-3, float(w0[0, 0].real), exit_wave=True)
    assert np.array_equal(img, got)
    assert wave.dtype == np.complex64 and orc.rel_l2(wave, x) < 1e-5
