"""Kirkland element tables and the 2-D electron scattering factors.

Restates the hot-path part of ``emsim/elem.py`` (reference):
``scattering_factor_function`` (``:120-156``, Kirkland 2010 eq. 5.17),
``scattering_factors2d`` (``:267-291``), ``elem_params`` (``:294-309``) and
``number`` (``:434-458``).  The parameter table is Kirkland's Appendix C fit,
read from ``data/kirkland_params.txt`` (one row per Z).

The reference stores the parameters as float32, evaluates the sum in float64
on a float64 frequency grid and stores the result as float32; this module
does the same, because the hip slice builder consumes the float32 table
exactly as the reference's fftw slice builder does.
"""
from functools import lru_cache
from pathlib import Path
from typing import List

import numpy as np

float_type = np.float32

_TABLE = Path(__file__).parent / "data" / "kirkland_params.txt"

_SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn "
    "Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
    "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn "
    "Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr").split()


@lru_cache(maxsize=None)
def _params_table() -> np.ndarray:
    """(103, 13) float32: chisq, a1..a3, b1..b3, c1..c3, d1..d3 per Z=1..103."""
    raw = np.loadtxt(_TABLE, dtype=np.float64, comments="#")
    if raw.shape != (103, 14) or not np.array_equal(raw[:, 0], np.arange(1, 104)):
        raise RuntimeError(f"{_TABLE} is malformed")
    tab = raw[:, 1:].astype(float_type)
    tab.setflags(write=False)
    return tab


def elem_params(elem_numbers: List[int]) -> np.ndarray:
    """[chisq, a1-3, b1-3, c1-3, d1-3] per Z as float32 (``elem.py:294-309``)."""
    if not all(1 <= int(z) <= 103 for z in elem_numbers):
        raise ValueError("z must be a interger in range [1, 92]")
    tab = _params_table()
    return np.array([tab[int(z) - 1] for z in elem_numbers], dtype=float_type)


_Z_OF = {s.upper(): z for z, s in enumerate(_SYMBOLS, start=1)}


def number(sym: str):
    """Element symbol -> Z, or None if unknown (``elem.py:434-458``).

    Matching is case-insensitive; the reference intends this (its table is
    keyed by upper-case symbols) and its own test asks ``number('He') == 2``.
    """
    return _Z_OF.get(str(sym).upper())


def symbol(z: int) -> str:
    return _SYMBOLS[int(z) - 1]


def scattering_factor_function(elem_numbers: List[int]):
    """f_e(q) = sum_k a_k/(q^2+b_k) + c_k exp(-d_k q^2), zero for q^2 >= 16.

    Follows ``elem.py:120-156``: float32 parameters, float64 arithmetic on the
    caller's grid, float32 result.
    """
    pms = elem_params(elem_numbers)
    n_elems = len(elem_numbers)

    def func_(q):
        q2 = q * q
        mask = q2 < 16
        out = np.empty(shape=(n_elems, *q.shape), dtype=float_type)
        for k in range(n_elems):
            a, b, cc, d = pms[k, 1:4], pms[k, 4:7], pms[k, 7:10], pms[k, 10:13]
            terms = [a[i] / (q2 + b[i]) + cc[i] * np.exp(-d[i] * q2) for i in range(3)]
            out[k] = np.sum(terms, axis=0)
            out[k] = np.where(mask, out[k], 0.)
        return out
    return func_


def scattering_factors2d(elem_numbers: List[int], n1: int, n2: int, d1: float, d2: float):
    """(nE, n1, n2) float32 form factors on the fftshifted grid (``elem.py:267-291``)."""
    kx = np.fft.fftshift(np.fft.fftfreq(n1, d1))
    ky = np.fft.fftshift(np.fft.fftfreq(n2, d2))
    kx, ky = np.meshgrid(kx, ky, indexing="ij")
    k = np.sqrt(kx * kx + ky * ky)
    return scattering_factor_function(elem_numbers)(k)


def scattering_factors_half(elem_numbers: List[int], n1: int, n2: int, d1: float, d2: float):
    """(nE, n1, n2//2+1) float32 table in FFT order, as the slice builders use it.

    Same steps as ``backend/slice_builder_base.py:30-33``: evaluate on the
    fftshifted grid, ifftshift back, keep the non-negative n2 frequencies.
    """
    sf = scattering_factors2d(list(elem_numbers), n1, n2, d1, d2)
    sf = np.fft.ifftshift(sf, axes=(-2, -1))
    return np.ascontiguousarray(sf[:, :, :n2 // 2 + 1], dtype=np.float32)
