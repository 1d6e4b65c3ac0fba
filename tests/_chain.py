"""Oracle versions of the reference's end-to-end chains, for the parity tests.

Each follows the reference's call sequence with the fftw backend
(emsim/pipe.py:46-84, emsim/pot.py:13-48), computed by oracle/.
"""
import numpy as np

from emsim_amd import atoms as atm
from emsim_amd import pot
from oracle import emsim_oracle as orc


def slices_oracle(mol, dz, pixel, roi=None, n_slices=None, bandlimit=True):
    """(V_fftw32 float64 as the reference returns it, V_f64, hist) for build_multi_slices."""
    mol, zs, counts = atm.sort_elements_and_count(mol)
    ns, n1, n2 = pot.finalize_slice_size(tuple(mol.space), (pixel, pixel), dz, roi, n_slices)
    hist = orc.bin_atoms_multi_slices(mol.coordinates, counts, ns, n1, n2, dz, pixel, pixel)
    sf = orc.scattering_factors_half(list(zs), n1, n2, pixel, pixel)
    return (orc.make_multi_slices_fftw32(hist, sf, bandlimit),
            orc.make_multi_slices_f64(hist, sf, bandlimit), hist)


def pipe_oracle(mol, tem, resolution, dz, roi):
    """(image_fftw32, image_f64, exit_fftw32, exit_f64) of Pipe(add_water=False).run."""
    pixel = 0.5 * resolution
    mol = atm.centralize(mol)
    v32, v64, _ = slices_oracle(mol, dz, pixel, roi)
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    n1, n2 = v32.shape[1:]
    w0 = orc.init_wave(n1, n2, pixel, pixel, tem.electron_dose)
    ex32 = orc.multislice_fftw32(w0, v32.astype(np.float32), dz, lam, gamma, pixel, pixel)
    ex64 = orc.multislice_f64(w0, v64, dz, lam, gamma, pixel, pixel)
    im32 = orc.intensity(orc.lens_propagate_fftw32(ex32, lam, tem.cs_mm, tem.defocus,
                                                   tem.aperture, pixel, pixel))
    im64 = np.abs(orc.lens_propagate_f64(ex64, lam, tem.cs_mm, tem.defocus, tem.aperture,
                                         pixel, pixel)) ** 2
    return im32, im64, ex32, ex64
