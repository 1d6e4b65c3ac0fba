"""Host-side logic of the API mirror (CPU only)."""
import math
import threading

import numpy as np
import pytest

import emsim_amd
from emsim_amd import atoms as atm
from emsim_amd import pot, synth
from emsim_amd.utils import rot
from emsim_amd.utils.array import assure_numpy


def test_sort_elements_and_count_with_must_include():
    mol = atm.AtomList(np.array([8, 6, 8, 1]), np.arange(12, dtype=np.float32).reshape(4, 3))
    s, zs, counts = atm.sort_elements_and_count(mol, [8, 16])
    assert list(zs) == [1, 6, 8, 16] and list(counts) == [1, 1, 2, 0]
    assert list(s.elements) == [1, 6, 8, 8]
    # rows move with their element
    np.testing.assert_array_equal(s.coordinates[0], [9, 10, 11])


def test_centralize_and_box_size():
    mol = atm.centralize(synth.five_atoms())
    np.testing.assert_allclose(mol.coordinates[:, 2], [-20, -10, 0, 10, 20])
    assert np.all(mol.coordinates[:, :2] == 0)
    # one-slice sizing: z extent 0 -> 3 A, thickness 5 A -> 1 slice
    assert pot.finalize_slice_size(tuple(mol.space), (50 / 512, 50 / 512), 5.0, (512, 512), 1) \
        == (1, 512, 512)
    assert atm.determine_box_size((0.0, 10.0, 10.5), (2.0, 1.0, 1.0), (None, None, None)) \
        == (2, 10, 11)
    assert atm.determine_box_size((90.0, 90.0, 90.0), 2.0, (None, 256, 256)) == (45, 256, 256)


def test_rotation_matrices_are_orthonormal_and_transposed_convention():
    q = rot.random_uniform_quaternions(20, rng=np.random.default_rng(0))
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, rtol=1e-12)
    r = rot.get_rotation_mattrices(q)
    np.testing.assert_allclose(r @ r.transpose(0, 2, 1), np.broadcast_to(np.eye(3), r.shape),
                               atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(r), 1.0, rtol=1e-12)
    # 90 degrees about z: the reference's matrix maps x -> -y (transpose of
    # the active rotation)
    qz = rot.get_quaternion([0, 0, 1], math.pi / 2)
    m = rot.get_rotation_mattrices(qz)[0]
    np.testing.assert_allclose(m @ [1, 0, 0], [0, -1, 0], atol=1e-12)
    # the global-RNG draw order is the reference's: 3 numbers per quaternion
    np.random.seed(5)
    a = rot.random_uniform_quaternion()
    np.random.seed(5)
    u1, u2, u3 = np.random.random(3)
    assert a[0] == np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)


def test_rotate_matches_matrix_product():
    mol = synth.protein_cloud(500, 1)
    q = rot.random_uniform_quaternion(np.random.default_rng(1))
    r = atm.rotate(mol, q)
    expect = (rot.get_rotation_mattrices(q)[0] @ mol.coordinates.astype(np.float64).T).T
    np.testing.assert_allclose(r.coordinates, expect.astype(np.float32), rtol=1e-6, atol=1e-5)


def test_protein_cloud_spacing_composition_and_seed():
    from scipy.spatial import cKDTree
    m = synth.protein_cloud(5000, 7)
    d, _ = cKDTree(m.coordinates).query(m.coordinates, k=2)
    assert d[:, 1].min() >= 1.0 - 1e-5
    counts = np.bincount(m.elements, minlength=17)[[1, 6, 7, 8, 16]]
    assert counts.sum() == 5000 and list(counts) == list(synth.element_counts(5000,
                                                                               synth.PROTEIN_FRACTIONS))
    np.testing.assert_array_equal(m.coordinates, synth.protein_cloud(5000, 7).coordinates)
    assert np.abs(m.coordinates).max() <= 0.5 * 90.0 * (5000 / 50000) ** (1 / 3) + 1e-4


def test_assure_numpy():
    import torch
    a = np.ones(3)
    assert assure_numpy(a) is a
    np.testing.assert_array_equal(assure_numpy(torch.ones(3)), np.ones(3))
    with pytest.raises(ValueError):
        assure_numpy([1, 2, 3])


class _FlakyPipe:
    """Stands in for Pipe: fails on chosen labels, returns a small image otherwise."""

    def __init__(self, fail_labels):
        self.fail_labels = set(fail_labels)
        self.thread = None

    def run(self, mol):
        self.thread = threading.current_thread().name
        if mol.label in self.fail_labels:
            raise RuntimeError(f"injected failure on {mol.label}")
        return np.full((2, 2), float(len(mol.elements)), np.float32)


def test_emsim_reports_errors_and_continues():
    mols = [atm.AtomList(np.ones(k, int), np.zeros((k, 3), np.float32), label=f"m{k}")
            for k in range(1, 6)]
    got = []
    pipe = _FlakyPipe({"m2", "m4"})
    sim = emsim_amd.simulator.EMSim(pipe, mols, lambda img, label: got.append((label, img[0, 0])))
    sim.run()
    assert got == [("m1", 1.0), ("m3", 3.0), ("m5", 5.0)]
    assert pipe.thread != threading.current_thread().name   # ran on the consumer thread


def test_emsim_iterator_that_raises_ends_the_stream():
    def gen():
        yield atm.AtomList(np.ones(1, int), np.zeros((1, 3), np.float32), label="a")
        raise IOError("source went away")
    got = []
    emsim_amd.simulator.EMSim(_FlakyPipe(()), gen(), lambda img, label: got.append(label)).run()
    assert got == ["a"]


def test_pipe_and_tem_repr_and_scherzer():
    tem = emsim_amd.tem.TEM(20, 200, 1.3, 700, 10e-3)
    tem.set_scherzer_condition()
    assert abs(tem.defocus - 699.319) < 1e-2
    p = emsim_amd.pipe.Pipe(tem, resolution=2.0, slice_thickness=2.0, roi=256)
    assert p.roi == (256, 256) and p.add_water is True and "ImagePipe" in repr(p)
    p.resolution = 1.0
    assert p._pixel_size == 0.5


def test_gpu_modules_import_without_a_gpu():
    """Every module of the package imports on a machine without a GPU (the
    device is touched only when work is launched)."""
    import importlib
    for name in ("emsim_amd.batch", "emsim_amd.learned", "emsim_amd.autotune",
                 "emsim_amd.train_fno", "emsim_amd.profiling", "emsim_amd.pipe",
                 "emsim_amd.simulator"):
        importlib.import_module(name)


def test_batched_groups_in_order():
    from emsim_amd.batch import batched
    assert list(batched(range(7), 3)) == [[0, 1, 2], [3, 4, 5], [6]]
    assert list(batched(range(6), 3)) == [[0, 1, 2], [3, 4, 5]]
    assert list(batched([], 4)) == []
    with pytest.raises(ValueError):
        list(batched(range(3), 0))


def test_pool_map_keeps_order_and_raises():
    from emsim_amd.batch import pool_map
    assert pool_map(lambda x: x * x, range(50), workers=8) == [x * x for x in range(50)]
    assert pool_map(lambda x: x + 1, [3], workers=8) == [4]
    assert pool_map(lambda x: x, [], workers=4) == []

    def bad(x):
        if x == 7:
            raise ValueError("seven")
        return x
    with pytest.raises(ValueError):
        pool_map(bad, range(10), workers=4)


def test_batched_quaternion_draws_equal_one_at_a_time():
    """random_uniform_quaternions draws all uniforms in one call; the values,
    and the generator state after the draw, equal n one-at-a-time Shoemake
    draws (the reference's order, rot.py:8-53), for a Generator and for
    numpy's global RNG."""
    from emsim_amd.utils.rot import random_uniform_quaternion, random_uniform_quaternions
    for seed in (0, 3, 20000):
        r1, r2 = np.random.default_rng(seed), np.random.default_rng(seed)
        one = np.array([random_uniform_quaternion(r1) for _ in range(257)])
        assert np.array_equal(one, random_uniform_quaternions(257, rng=r2))
        assert r1.random() == r2.random()
        np.random.seed(seed)
        one = np.array([random_uniform_quaternion() for _ in range(9)])
        np.random.seed(seed)
        assert np.array_equal(one, random_uniform_quaternions(9))
    q = random_uniform_quaternions(1000, rng=np.random.default_rng(1))
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0)
