"""Synthetic specimens of the large-slab and mixed-size configurations (CPU)."""
import numpy as np

from emsim_amd import atoms as atm
from emsim_amd import synth


def test_water_geometry_density_and_isotropy():
    w = synth.water_slab(8.0, 80.0, 80.0, density=0.0334, seed=3)
    n_o = int(round(0.0334 * 8 * 80 * 80))
    assert len(w) == 3 * n_o
    assert np.array_equal(np.bincount(w.elements, minlength=9)[[1, 8]], [2 * n_o, n_o])
    c = w.coordinates.astype(np.float64).reshape(-1, 3, 3)
    b1, b2 = c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]
    d1, d2 = np.linalg.norm(b1, axis=1), np.linalg.norm(b2, axis=1)
    assert np.allclose(d1, 0.96, atol=1e-4) and np.allclose(d2, 0.96, atol=1e-4)
    ang = np.degrees(np.arccos(np.sum(b1 * b2, axis=1) / (d1 * d2)))
    assert np.allclose(ang, 104.5, atol=0.02)
    o = c[:, 0]
    assert (o >= [-4, -40, -40]).all() and (o < [4, 40, 40]).all()
    bis = b1 + b2
    bis /= np.linalg.norm(bis, axis=1)[:, None]
    assert np.abs(bis.mean(axis=0)).max() < 5 / np.sqrt(n_o)      # random orientations
    assert np.array_equal(synth.water_slab(8.0, 80.0, 80.0, seed=3).coordinates, w.coordinates)


def test_large_slab_composition():
    s = synth.large_slab(lz=20.0, lx=200.0, ly=200.0, n_clouds=2, cloud_atoms=3000, seed=1)
    n_o = int(round(0.0334 * 20 * 200 * 200))
    assert len(s) == 3 * n_o + 2 * 3000
    assert set(np.unique(s.elements)) <= {1, 6, 7, 8, 16}


def test_mixed_workload_sizes_and_counts():
    items = synth.mixed_workload(count=12, seed=2, lo=500, hi=2000)
    assert len(items) == 12
    assert {n for _, n in items} <= {128, 256, 512}
    for mol, _ in items:
        assert 500 <= len(mol) <= 2000
        assert np.isfinite(mol.coordinates).all()
    # deterministic
    again = synth.mixed_workload(count=12, seed=2, lo=500, hi=2000)
    assert all(np.array_equal(a.coordinates, b.coordinates) and n == m
               for (a, n), (b, m) in zip(items, again))
