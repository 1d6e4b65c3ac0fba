"""Synthetic specimens for benchmarks and tests (no PDB files, no network).

``protein_cloud`` draws a protein-like atom cloud: uniform positions in a
sphere, a minimum spacing enforced by rejection, and the element mix of a
typical protein (H 0.49 / C 0.32 / N 0.09 / O 0.095 / S 0.005).  The draw is
a pure function of the seed (numpy PCG64).  ``five_atoms`` is the
Kirkland (2010) p.95 example used by the reference's demo
(``demos/five_atoms.py:19-24``).
"""
from typing import Sequence

import numpy as np

from .atoms import AtomList
from .utils.rot import get_rotation_mattrices

PROTEIN_ZS = (1, 6, 7, 8, 16)
PROTEIN_FRACTIONS = (0.49, 0.32, 0.09, 0.095, 0.005)


def element_counts(n_atoms: int, fractions: Sequence[float]) -> np.ndarray:
    """Integer counts summing to n_atoms, largest-remainder rounding."""
    raw = np.asarray(fractions, dtype=np.float64) * n_atoms
    counts = np.floor(raw).astype(np.int64)
    short = n_atoms - int(counts.sum())
    order = np.argsort(-(raw - counts), kind="stable")
    counts[order[:short]] += 1
    return counts


def _sphere_points(rng, n, radius):
    pts = rng.uniform(-radius, radius, size=(int(n * 2.2) + 64, 3))
    pts = pts[np.einsum("ij,ij->i", pts, pts) <= radius * radius]
    return pts


def protein_cloud(n_atoms: int, seed: int, diameter: float = None, min_spacing: float = 1.0,
                  label: str = "") -> AtomList:
    """Uniform cloud of n_atoms in a sphere with pairwise spacing >= min_spacing.

    diameter defaults to 90 A at 50k atoms, scaled as n_atoms^(1/3) (constant
    density).  Candidates are drawn in rounds and a candidate closer than
    min_spacing to an accepted atom, or to an earlier candidate of its own
    round, is rejected (a cell grid of side min_spacing finds neighbours).
    """
    from scipy.spatial import cKDTree

    if diameter is None:
        diameter = 90.0 * (n_atoms / 50000.0) ** (1.0 / 3.0)
    radius = 0.5 * diameter
    rng = np.random.default_rng(seed)
    accepted = np.empty((0, 3))
    for _ in range(200):
        need = n_atoms - accepted.shape[0]
        if need <= 0:
            break
        cand = _sphere_points(rng, max(need, 256), radius)
        tree_c = cKDTree(cand)
        drop = np.zeros(cand.shape[0], dtype=bool)
        pairs = tree_c.query_pairs(r=min_spacing, output_type="ndarray")
        if pairs.size:
            drop[np.maximum(pairs[:, 0], pairs[:, 1])] = True
        if accepted.shape[0]:
            d, _ = cKDTree(accepted).query(cand, k=1, distance_upper_bound=min_spacing)
            drop |= d < min_spacing
        keep = cand[~drop][:need]
        accepted = np.concatenate([accepted, keep])
    if accepted.shape[0] < n_atoms:
        raise RuntimeError(f"could not place {n_atoms} atoms at spacing {min_spacing} "
                           f"in a {diameter:.1f} A sphere")
    counts = element_counts(n_atoms, PROTEIN_FRACTIONS)
    zs = np.repeat(np.asarray(PROTEIN_ZS, dtype=np.int64), counts)
    zs = zs[rng.permutation(n_atoms)]
    return AtomList(elements=zs, coordinates=accepted.astype(np.float32), label=label)


def rotated(mol: AtomList, quat: np.ndarray) -> AtomList:
    """Rotate about the bounding-box centre with the reference's matrix convention."""
    c = mol.coordinates
    centre = c.min(axis=0) + (c.max(axis=0) - c.min(axis=0)) / 2.0
    rot = get_rotation_mattrices(quat)[0]
    out = ((c - centre).astype(np.float64) @ rot.T).astype(np.float32)
    return AtomList(elements=mol.elements, coordinates=out, label=mol.label)


def five_atoms() -> AtomList:
    """C, Si, Cu, Au, U on the z=0 plane at x=0, y=5..45 A (not yet centred)."""
    return AtomList(elements=np.array([6, 14, 29, 79, 92], dtype=int),
                    coordinates=np.array([[0, 0, 5], [0, 0, 15], [0, 0, 25], [0, 0, 35],
                                          [0, 0, 45]], dtype=np.float32))
