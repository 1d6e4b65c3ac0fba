"""Shared propagator state (``emsim/backend/wave_propagator_base.py:10-29``).

The wavelength and relativistic gamma are float64 host constants; the hip
kernels receive them rounded to float32, as the reference's fftw binding
does (``emsim/backend/fftw_ext/src/wave_kernel.cpp:21``).
"""
import numpy as np

from ..physics import electron_relativity_gamma, electron_wave_length_angstrom


class WavePropagatorBase:
    def __init__(self, n1: int, n2: int, d1: float, d2: float, beam_energy_kev: float):
        self.wave_shape = (int(n1), int(n2))
        self.d1, self.d2 = float(d1), float(d2)
        self.wave_length = electron_wave_length_angstrom(beam_energy_kev)
        self.relativity_gamma = electron_relativity_gamma(beam_energy_kev)

    def init_wave_value(self, electron_dose: float) -> np.float32:
        """sqrt(dose d1 d2) rounded to float32 (``wave_propagator_base.py:25-29``)."""
        return np.float32(np.sqrt(electron_dose * self.d1 * self.d2))

    def init_wave(self, electron_dose: float):
        raise NotImplementedError

    def slice_transmit(self, wave, aslice):
        raise NotImplementedError

    def space_propagate(self, wave, dz):
        raise NotImplementedError

    def multislice_propagate(self, wave, slices, dz: float):
        raise NotImplementedError

    def singleslice_propagate(self, wave, aslice, dz: float):
        raise NotImplementedError

    def lens_propagate(self, wave, cs_mm: float, defocus: float, aperture: float):
        raise NotImplementedError
