"""Atom lists: sorting by element, centring, rotation, box sizing.

Restates the hot-path part of ``emsim/atoms.py`` (reference): ``AtomList``
(``:21-50``), ``sort_elements_and_count`` (``:53-86``), ``concatenate``
(``:117-120``), ``translate`` / ``centralize`` (``:123-159``), ``rotate`` /
``orientations_gen`` (``:162-211``) and ``determine_box_size``
(``:214-223``).  Coordinates are float32 with columns (z, x, y): column 0 is
the slice axis, columns 1 and 2 the two lateral axes (n1, n2).
"""
import bisect
import math
from typing import Generator, List, Optional, Tuple, Union

import numpy as np

from .utils.rot import get_rotation_mattrices

float_type = np.float32


class AtomList(object):
    def __init__(self, elements: np.ndarray, coordinates: np.ndarray, label=""):
        self.elements = elements
        self.coordinates = coordinates.astype(float_type)
        self.label = label

    @property
    def r_min(self):
        return self.coordinates.min(axis=0)

    @property
    def r_max(self):
        return self.coordinates.max(axis=0)

    @property
    def space(self):
        return self.r_max - self.r_min

    def __len__(self):
        return len(self.elements)

    def __repr__(self):
        return f"<AtomList consists of {len(self.elements)} atoms, label={self.label}>"


def sort_elements_and_count(atom_list: AtomList, must_include_elems: Optional[List[int]] = ()):
    """Sort atoms by Z; return (sorted list, unique Z, counts).

    Elements in ``must_include_elems`` that do not occur are inserted with a
    zero count at their sorted position, as ``atoms.py:76-84`` does.
    """
    idx = np.argsort(atom_list.elements)
    elements = atom_list.elements[idx]
    coordinates = atom_list.coordinates[idx]
    unique_elements, counts = np.unique(elements, return_counts=True)
    for key_z in sorted(list(must_include_elems)):
        if key_z not in unique_elements:
            i = bisect.bisect_left(unique_elements, key_z)
            unique_elements = np.insert(unique_elements, i, key_z)
            counts = np.insert(counts, i, 0)
    return (AtomList(elements=elements, coordinates=coordinates, label=atom_list.label),
            unique_elements, counts)


def concatenate(atmls: List[AtomList]) -> AtomList:
    return AtomList(elements=np.concatenate([a.elements for a in atmls]),
                    coordinates=np.concatenate([a.coordinates for a in atmls]))


def translate(atom_list: AtomList, r) -> AtomList:
    return AtomList(elements=atom_list.elements, coordinates=atom_list.coordinates + r,
                    label=atom_list.label)


def centralize(atom_list: AtomList) -> AtomList:
    """Move the bounding-box centre to the origin: r - r_min - space/2."""
    return translate(atom_list, -atom_list.r_min - atom_list.space / 2.0)


def rotate(atom_list: AtomList, quat: np.ndarray, set_center: bool = False) -> AtomList:
    """Apply the reference's quaternion matrix to every atom (float64 matmul)."""
    if set_center:
        atom_list = centralize(atom_list)
    rot = get_rotation_mattrices(quat)
    rot_coords = np.matmul(rot, atom_list.coordinates[..., None]).squeeze()
    return AtomList(elements=atom_list.elements, coordinates=rot_coords, label=atom_list.label)


def orientations_gen(atom_list: AtomList, quats: np.ndarray, set_center: bool = False) \
        -> Generator[AtomList, None, None]:
    for i in range(quats.shape[0]):
        yield rotate(atom_list, quats[i], set_center)


def determine_box_size(space: Tuple[float, float, float],
                       voxel_size: Union[float, Tuple[float, float, float]],
                       box_size: Tuple[Optional[int], Optional[int], Optional[int]]):
    """ceil(space / voxel) per axis; extents below 1e-3 A count as 3 A."""
    space = [3.0 if space[d] < 1e-3 else space[d] for d in range(3)]
    if type(voxel_size) is float:
        voxel_size = (voxel_size,) * 3
    box = [int(math.ceil(length / d)) for length, d in zip(space, voxel_size)]
    return tuple(box[d] if box_size[d] is None else box_size[d] for d in range(3))
