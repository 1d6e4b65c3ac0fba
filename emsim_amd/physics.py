"""Electron-optics constants and closed forms (host side, float64).

Restates ``emsim/physics.py`` (reference): CGS constants (``:5-20``),
relativistic gamma (``:23-25``), wavelength (``:44-59``), interaction
parameter (``:62-81``), aberration / MTF / CTF (``:85-111``) and the
Scherzer condition (``:114-156``).  These are the only float64 numbers the
hip backend uses; they are rounded to float32 before reaching a kernel.
"""
import math

import numpy as np

# CGS constants, as in the reference
water_num_dens = 0.031            # water molecules per A^3
e = 4.8032042510e-10              # statC
h = 6.62607015e-27                # erg s
hbar = h / 2.0 / math.pi
m0 = 9.1093835611e-28             # g
a0 = 5.29177210903e-9             # cm
c = 29979245800.0                 # cm / s
m0c2 = m0 * c ** 2                # erg
hc = h * c
hbar_c = hbar * c

e_coulumb = 1.602176634e-19
m0c2_keV = m0c2 * 1e-7 / e_coulumb / 1000.
hbar_c_keV_Ang = hbar_c * 0.01 / e_coulumb
hc_keV_Ang = hc * 0.01 / e_coulumb


def electron_relativity_gamma(beam_energy_kev: float) -> float:
    """gamma = 1 + E / m0c^2 (``physics.py:23-25``)."""
    return 1 + beam_energy_kev / m0c2_keV


def electron_relativity_mass(beam_energy_kev: float) -> float:
    """Relativistic rest energy gamma * m0 c^2 in erg (``physics.py:28-41``)."""
    return electron_relativity_gamma(beam_energy_kev) * m0c2


def electron_wave_length_angstrom(beam_energy_kev: float) -> float:
    """lambda = hc / sqrt(E^2 + 2 E m0c^2), in Angstrom (``physics.py:44-59``)."""
    return hc_keV_Ang / math.sqrt(beam_energy_kev ** 2 + 2 * m0c2_keV * beam_energy_kev)


def interaction_parameter(beam_energy_kev: float) -> float:
    """sigma in 1/(E*L) units (``physics.py:62-81``)."""
    dimensionless = (m0c2_keV + beam_energy_kev) / math.sqrt(
        beam_energy_kev * (2 * m0c2_keV + beam_energy_kev))
    return (e / hbar_c) * dimensionless


def aberration(wave_length_angstrom: float, cs_mm: float, defocus_angstrom: float):
    """chi(k) = pi (K^4/2 - D K^2) in reduced units (``physics.py:85-92``)."""
    cslambda = cs_mm * 1e7 * wave_length_angstrom
    big_d = defocus_angstrom / math.sqrt(cslambda)

    def aberration_(k):
        big_k = k * pow(cslambda * wave_length_angstrom ** 2, 0.25)
        return math.pi * (0.5 * big_k ** 4 - big_d * big_k ** 2)
    return aberration_


def mtf(wave_length_angstrom: float, cs_mm: float, defocus_angstrom: float):
    """exp(-i chi(k)) (``physics.py:95-102``)."""
    def mtf_(k):
        return np.exp(-1j * aberration(wave_length_angstrom, cs_mm, defocus_angstrom)(k))
    return mtf_


def ctf(wave_length_angstrom: float, cs_mm: float, defocus_angstrom: float):
    """sin(chi(k)) (``physics.py:105-111``)."""
    def ctf_(k):
        return np.sin(aberration(wave_length_angstrom, cs_mm, defocus_angstrom)(k))
    return ctf_


def optimal_focus(wave_length_angstrom: float, cs_mm: float, nd: int) -> float:
    """df = sqrt((2 nd - 1/2) Cs lambda) (``physics.py:114-134``)."""
    cslambda = cs_mm * 1e7 * wave_length_angstrom
    return math.sqrt((2 * nd - 0.5) * cslambda)


def scherzer_condition(wave_length_angstrom: float, cs_mm: float):
    """(Scherzer defocus [A], Scherzer aperture [rad]) (``physics.py:136-156``)."""
    df = optimal_focus(wave_length_angstrom, cs_mm, 1)
    aper = pow(6 * wave_length_angstrom / (cs_mm * 1e7), 0.25)
    return df, aper
