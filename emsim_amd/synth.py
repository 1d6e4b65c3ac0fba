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


def water_slab(lz: float, lx: float, ly: float, density: float = 0.0334, seed: int = 1,
               oh: float = 0.96, hoh_deg: float = 104.5, chunk: int = 1 << 21) -> AtomList:
    """Amorphous water: oxygens uniform in the box [-lz/2, lz/2) x [-lx/2,
    lx/2) x [-ly/2, ly/2) at `density` per A^3, two hydrogens per oxygen at
    `oh` A and `hoh_deg` degrees in a uniformly random orientation.  Atoms
    are ordered O, H, H per molecule; coordinates (z, x, y)."""
    from .utils.rot import random_uniform_quaternions
    rng = np.random.default_rng(seed)
    n_o = int(round(density * lz * lx * ly))
    half = 0.5 * np.deg2rad(hoh_deg)
    # the two O-H bonds in the molecule frame
    bonds = np.array([[np.cos(half), np.sin(half), 0.0], [np.cos(half), -np.sin(half), 0.0]]) * oh
    out = np.empty((3 * n_o, 3), dtype=np.float32)
    lo = np.array([-lz / 2, -lx / 2, -ly / 2])
    hi = -lo
    for s in range(0, n_o, chunk):
        m = min(chunk, n_o - s)
        o = rng.uniform(lo, hi, size=(m, 3))
        rot = get_rotation_mattrices(random_uniform_quaternions(m, rng=rng))   # (m, 3, 3)
        h = np.einsum("mij,bj->mbi", rot, bonds)                              # (m, 2, 3)
        blk = out[3 * s:3 * (s + m)].reshape(m, 3, 3)
        blk[:, 0] = o
        blk[:, 1] = o + h[:, 0]
        blk[:, 2] = o + h[:, 1]
    elements = np.tile(np.array([8, 1, 1], dtype=np.int64), n_o)
    return AtomList(elements=elements, coordinates=out, label=f"water {lz}x{lx}x{ly}")


def large_slab(lz: float = 400.0, lx: float = 1024.0, ly: float = 1024.0, n_clouds: int = 5,
               cloud_atoms: int = 50000, seed: int = 1) -> AtomList:
    """Water slab plus n_clouds protein-like clouds at random positions
    inside it (the large-slab configuration: 2048^2 waves of 0.5 A pixels,
    400 slices of 1 A)."""
    from .atoms import concatenate
    slab = water_slab(lz, lx, ly, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    parts = [slab]
    for i in range(n_clouds):
        c = protein_cloud(cloud_atoms, seed + 2000 + i)
        d = 90.0 * (cloud_atoms / 50000.0) ** (1.0 / 3.0)
        room = np.maximum(np.array([lz, lx, ly]) / 2 - d / 2, 0.0)
        shift = rng.uniform(-room, room).astype(np.float32)
        parts.append(AtomList(c.elements, c.coordinates + shift))
    return concatenate(parts)


def mixed_workload(count: int = 256, seed: int = 2, lo: int = 5000, hi: int = 200000,
                   sizes=(128, 256, 512)):
    """The batched mixed-size configuration: `count` clouds with atom counts
    log-uniform in [lo, hi] (radius scaled as N^(1/3)), uniform random
    orientations (quaternion seed `seed`) and wave sizes drawn from `sizes`.
    Returns [(AtomList, wave size)]."""
    from .utils.rot import random_uniform_quaternions
    rng = np.random.default_rng(seed + 7)
    n_atoms = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(int)
    wave = rng.choice(np.asarray(sizes), size=count)
    quats = random_uniform_quaternions(count, rng=np.random.default_rng(seed))
    from .batch import pool_map   # clouds are independent: build them on a thread pool
    mols = pool_map(lambda i: rotated(protein_cloud(int(n_atoms[i]), seed + 100 + i), quats[i]),
                    range(count))
    return [(m, int(w)) for m, w in zip(mols, wave)]


def five_atoms() -> AtomList:
    """C, Si, Cu, Au, U on the z=0 plane at x=0, y=5..45 A (not yet centred)."""
    return AtomList(elements=np.array([6, 14, 29, 79, 92], dtype=int),
                    coordinates=np.array([[0, 0, 5], [0, 0, 15], [0, 0, 25], [0, 0, 35],
                                          [0, 0, 45]], dtype=np.float32))
