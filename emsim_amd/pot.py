"""Projected-potential frontends (``emsim/pot.py:13-74``).

``build_one_slice`` and ``build_multi_slices`` keep the reference's
signatures and steps: sort atoms by element, size the box, make a builder
from the current backend, bin, optionally add water, build the slices.
"""
from typing import Optional, Tuple, Union

import numpy as np

from . import atoms as atm
from . import config

float_type = np.float32


def build_one_slice(mol: atm.AtomList, pixel_size: Union[float, Tuple[float, float]],
                    lateral_size: Optional[Union[int, Tuple[int, int]]] = None):
    one_slice_builder = config.get_current_backend().one_slice_builder
    if type(pixel_size) is float:
        pixel_size = (pixel_size,) * 2
    mol, unique_elements, counts = atm.sort_elements_and_count(mol)
    _, n1, n2 = finalize_slice_size(tuple(mol.space), pixel_size, 5.0, lateral_size, 1)
    builder = one_slice_builder(unique_elements, n1, n2, pixel_size[0], pixel_size[1])
    atmv = builder.bin_atoms_one_slice(mol.coordinates, counts)
    return builder.make_one_slice(atmv)


def build_multi_slices(mol: atm.AtomList, dz: float,
                       pixel_size: Union[float, Tuple[float, float]],
                       lateral_size: Optional[Union[int, Tuple[int, int]]] = None,
                       n_slices: Optional[int] = None, add_water: bool = False):
    if type(pixel_size) is float:
        pixel_size = (pixel_size,) * 2
    multi_slices_builder = config.get_current_backend().multi_slices_builder
    must_include_elems = [1, 8] if add_water else []
    mol, unique_elements, counts = atm.sort_elements_and_count(mol, must_include_elems)
    n_slices, n1, n2 = finalize_slice_size(tuple(mol.space), pixel_size, dz, lateral_size,
                                           n_slices)
    builder = multi_slices_builder(unique_elements, n_slices, n1, n2, dz, pixel_size[0],
                                   pixel_size[1])
    atmv = builder.bin_atoms_multi_slices(mol.coordinates, counts)
    if add_water:
        atmv = builder.add_water(atmv)
    return builder.make_multi_slices(atmv)


def finalize_slice_size(mol_space, pixel_size, thickness: float,
                        lateral_size: Optional[Union[int, Tuple[int, int]]] = None,
                        n_slices: Optional[int] = None):
    dims = [None, None, None]
    if lateral_size is not None:
        if isinstance(lateral_size, int):
            dims[1] = dims[2] = lateral_size
        elif isinstance(lateral_size, tuple) and len(lateral_size) == 2:
            dims[1], dims[2] = lateral_size
    if isinstance(n_slices, int):
        dims[0] = n_slices
    if isinstance(pixel_size, float):
        pixel_size = (pixel_size, pixel_size)
    return atm.determine_box_size(tuple(mol_space),
                                  voxel_size=(thickness, pixel_size[0], pixel_size[1]),
                                  box_size=(dims[0], dims[1], dims[2]))
