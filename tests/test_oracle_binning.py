"""Pins of the binning oracle and of the kernel's float32 edge algorithm (CPU).

The reference bins with numpy.histogramdd on float64 numpy.arange edges
(emsim/backend/slice_builder_base.py:180-225).  The oracle calls numpy the
same way; these tests pin it against a per-atom brute force, and pin the
HIP kernel's float32 edge-table algorithm (emulated here in numpy) against
the same brute force.  The library's host edge builder is compared with
numpy.arange bit for bit.
"""
import numpy as np
import pytest

from oracle import emsim_oracle as orc

F32 = np.float32


def kernel_bin_of(x: np.float32, e32: np.ndarray, last_exact: bool, inv_d: np.float32) -> int:
    """numpy emulation of bin_of() in emsim_amd/csrc/kernels_slice.hpp."""
    nb = len(e32) - 1
    x = F32(x)
    if not (x >= e32[0]):
        return -1
    if x >= e32[nb]:
        return nb - 1 if (x == e32[nb] and last_exact) else -1
    b = int(np.floor(F32(F32(x - e32[0]) * inv_d)))
    b = min(max(b, 0), nb - 1)
    while b + 1 < nb and e32[b + 1] <= x:
        b += 1
    while b > 0 and e32[b] > x:
        b -= 1
    return b


def round_up_f32(e64: np.ndarray) -> np.ndarray:
    f = e64.astype(F32)
    low = f.astype(np.float64) < e64
    f[low] = np.nextafter(f[low], F32(np.inf))
    return f


def edge_cases(edges):
    """Coordinates on, just below and just above every edge, plus specials."""
    e32 = edges.astype(F32)
    pts = [e32, np.nextafter(e32, F32(-np.inf)), np.nextafter(e32, F32(np.inf)),
           (edges[:-1] + 0.5 * np.diff(edges)).astype(F32),
           np.array([np.nan, np.inf, -np.inf, 0.0, -0.0], dtype=F32)]
    return np.concatenate(pts)


@pytest.mark.parametrize("n,d", [(8, 1.0), (7, 1.0), (5, 2.0), (16, 0.5), (45, 2.0),
                                 (512, 50 / 512), (33, 0.3), (10, 0.1), (64, 1.7), (100, 1.0)])
def test_oracle_edges_are_reference_arange(n, d):
    e = orc.bin_edges(n, d)
    start = -d * (n // 2)
    np.testing.assert_array_equal(e, np.arange(start, start + n * d + d, d))


@pytest.mark.parametrize("n,d", [(8, 1.0), (7, 1.0), (45, 2.0), (512, 50 / 512), (33, 0.3),
                                 (10, 0.1), (64, 1.7), (2048, 0.5)])
def test_kernel_edge_algorithm_matches_bruteforce(n, d):
    edges = orc.bin_edges(n, d)
    e32 = round_up_f32(edges)
    last_exact = float(F32(edges[-1])) == float(edges[-1])
    inv_d = F32(1.0 / d)
    xs = edge_cases(edges)
    for x in xs:
        assert kernel_bin_of(x, e32, last_exact, inv_d) == orc.bin_index_bruteforce(x, edges), x


def test_bruteforce_matches_histogramdd_on_edge_cases():
    rng = np.random.default_rng(3)
    shape, steps = (5, 6, 7), (2.0, 0.3, 1.0)
    edges = [orc.bin_edges(n, d) for n, d in zip(shape, steps)]
    cols = [rng.choice(edge_cases(e), size=400) for e in edges]
    pts = np.stack(cols, axis=1).astype(F32)
    pts = pts[np.isfinite(pts).all(1)]
    h_np = orc.bin_atoms_multi_slices(pts, [len(pts)], *shape, *steps)[0]
    h_bf = np.zeros([len(e) - 1 for e in edges], dtype=F32)
    for p in pts:
        idx = [orc.bin_index_bruteforce(p[a], edges[a]) for a in range(3)]
        if min(idx) >= 0:
            h_bf[tuple(idx)] += 1
    np.testing.assert_array_equal(h_np, h_bf)


def test_multi_slice_binning_uses_elements_in_order():
    coords = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [-1.5, 1.2, -0.7], [9e9, 0, 0]], F32)
    h = orc.bin_atoms_multi_slices(coords, [2, 2], 4, 4, 4, 1.0, 1.0, 1.0)
    assert h.shape == (2, 4, 4, 4) and h.dtype == F32
    assert h[0, 2, 2, 2] == 2 and h[0].sum() == 2
    assert h[1, 0, 3, 1] == 1 and h[1].sum() == 1     # the far atom is dropped


def test_one_slice_binning_ignores_z():
    coords = np.array([[100.0, 0.0, 0.0], [-100.0, 0.0, 0.0]], F32)
    h = orc.bin_atoms_one_slice(coords, [2], 4, 4, 1.0, 1.0)
    assert h[0, 2, 2] == 2


def test_nonuniform_arange_length_is_detected():
    # numpy.arange may return n + 2 edges for non-dyadic steps; the reference
    # then builds an (n+1)-bin axis.  Find such a case to make sure it exists.
    found = False
    for n in range(3, 200):
        for d in (0.1, 0.3, 0.7, 1.1):
            if len(orc.bin_edges(n, d)) != n + 1:
                found = True
                break
        if found:
            break
    assert found


def test_library_edges_equal_numpy_arange():
    hip_ext = pytest.importorskip("emsim_amd.backend.hip_ext")
    rng = np.random.default_rng(7)
    cases = [(n, d) for n in (1, 2, 7, 8, 45, 100, 256, 512, 2048)
             for d in (1.0, 2.0, 0.5, 50 / 512, 0.3, 0.1, 1.7)]
    cases += [(int(rng.integers(1, 3000)), float(rng.uniform(0.05, 4.0))) for _ in range(300)]
    for n, d in cases:
        e64, e32 = hip_ext.bin_edges(n, d)
        np.testing.assert_array_equal(e64, orc.bin_edges(n, d), err_msg=f"n={n} d={d}")
        np.testing.assert_array_equal(e32, round_up_f32(e64))
        assert np.all(e32.astype(np.float64) >= e64)
        assert np.all(np.nextafter(e32, F32(-np.inf)).astype(np.float64) < e64)
