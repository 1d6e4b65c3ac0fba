"""Shared state of the slice builders (``emsim/backend/slice_builder_base.py``).

Holds the grid, the sorted element list and the (nE, n1, n2//2+1) float32
form-factor table in FFT order (``slice_builder_base.py:30-33``).  The table
is a host-side constant; it is cached per (elements, grid) because the
reference's ``Pipe`` builds a new builder for every image.
"""
from functools import lru_cache
from typing import List

import numpy as np

from .. import elem


@lru_cache(maxsize=64)
def _cached_table(zs: tuple, n1: int, n2: int, d1: float, d2: float) -> np.ndarray:
    tab = elem.scattering_factors_half(list(zs), n1, n2, d1, d2)
    tab.setflags(write=False)
    return tab


def scattering_factor_table(unique_elements, n1, n2, d1, d2) -> np.ndarray:
    zs = tuple(int(z) for z in sorted(unique_elements))
    return _cached_table(zs, int(n1), int(n2), float(d1), float(d2))


class OneSliceBuilderBase:
    def __init__(self, unique_elements: List[int], n1: int, n2: int, d1: float, d2: float):
        self.n1, self.n2 = int(n1), int(n2)
        self.d1, self.d2 = float(d1), float(d2)
        self.scattering_factors = scattering_factor_table(unique_elements, n1, n2, d1, d2)
        self.unique_elements = unique_elements

    def bin_atoms_one_slice(self, atom_coordinates_sorted_by_elems, unique_elements_count):
        raise NotImplementedError

    def make_one_slice(self, atom_histograms_one_slice, symmetric_bandlimit: bool = True):
        raise NotImplementedError


class MultiSlicesBuilderBase:
    def __init__(self, unique_elements: List[int], n_slices: int, n1: int, n2: int,
                 dz: float, d1: float, d2: float):
        self.n_slice = int(n_slices)
        self.n1, self.n2 = int(n1), int(n2)
        self.dz = float(dz)
        self.d1, self.d2 = float(d1), float(d2)
        self.scattering_factors = scattering_factor_table(unique_elements, n1, n2, d1, d2)
        self.unique_elements = unique_elements

    def bin_atoms_multi_slices(self, atom_coordinates_sorted_by_elems, unique_elements_count):
        raise NotImplementedError

    def add_water(self, atom_histograms):
        raise NotImplementedError

    def make_multi_slices(self, atom_histograms, symmetric_bandlimit: bool = True):
        raise NotImplementedError
