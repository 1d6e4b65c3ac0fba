"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Bars (stated per test): binning bit-exact; potentials, waves and images
within 1e-5 relative L2 of the fftw-semantics oracle; the distance to the
FP64 oracle is asserted at a looser bound and printed.
"""
import numpy as np
import pytest
import torch

import emsim_amd
from emsim_amd import atoms as atm
from emsim_amd import pot, synth
from oracle import emsim_oracle as orc
from tests._chain import pipe_oracle, slices_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5          # relative L2 vs the fftw-semantics oracle (the parity bar)
TOL64 = 1e-5        # relative L2 vs the FP64 oracle
LAM, GAMMA = orc.wavelength_and_gamma(200)


def host(t):
    return t.detach().cpu().numpy()


def report(name, got, ref32, ref64):
    e32, e64 = orc.rel_l2(got, ref32), orc.rel_l2(got, ref64)
    print(f"[{name}] rel L2 vs fftw32 oracle {e32:.3e}, vs FP64 oracle {e64:.3e}", flush=True)
    return e32, e64


@pytest.fixture(autouse=True)
def _backend(hip):
    emsim_amd.config.set_backend("hip")
    yield
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# binning
# ---------------------------------------------------------------------------
def _edge_cloud(shape, steps, n, seed):
    rng = np.random.default_rng(seed)
    cols = []
    for nn, d in zip(shape, steps):
        e = orc.bin_edges(nn, d).astype(np.float32)
        pool = np.concatenate([e, np.nextafter(e, np.float32(-np.inf)),
                               np.nextafter(e, np.float32(np.inf)),
                               rng.uniform(e[0] - 2 * d, e[-1] + 2 * d, 64).astype(np.float32)])
        cols.append(rng.choice(pool, size=n))
    pts = np.stack(cols, axis=1).astype(np.float32)
    pts[::97, 1] = np.nan
    return pts


@pytest.mark.parametrize("shape,steps", [((45, 256, 256), (2.0, 1.0, 1.0)),
                                         ((7, 64, 32), (0.25, 0.7, 50 / 512)),
                                         ((1, 16, 16), (5.0, 1.0, 1.0))])
def test_bin_atoms_bit_exact_edge_cases(shape, steps):
    pts = _edge_cloud(shape, steps, 20000, 11)
    counts = np.array([5000, 0, 9000, 6000], dtype=np.int64)
    from emsim_amd.backend.slice_builder_hip import MultiSlicesBuilder
    b = MultiSlicesBuilder([1, 6, 7, 8], *shape, steps[0], steps[1], steps[2])
    got = host(b.bin_atoms_multi_slices(pts, counts))
    ref = orc.bin_atoms_multi_slices(pts, counts, *shape, *steps)
    assert got.dtype == np.float32 and got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


def test_bin_atoms_bit_exact_protein_cloud():
    mol = atm.centralize(synth.protein_cloud(50000, 0))
    mol, zs, counts = atm.sort_elements_and_count(mol)
    from emsim_amd.backend.slice_builder_hip import MultiSlicesBuilder
    b = MultiSlicesBuilder(zs, 45, 256, 256, 2.0, 1.0, 1.0)
    got = host(b.bin_atoms_multi_slices(mol.coordinates, counts))
    ref = orc.bin_atoms_multi_slices(mol.coordinates, counts, 45, 256, 256, 2.0, 1.0, 1.0)
    np.testing.assert_array_equal(got, ref)
    assert got.sum() == ref.sum() > 0


def test_bin_one_slice_bit_exact():
    pts = _edge_cloud((1, 128, 128), (1.0, 0.5, 0.5), 5000, 2)
    pts[:, 0] = np.random.default_rng(0).uniform(-1e3, 1e3, len(pts))
    counts = np.array([2000, 3000], dtype=np.int64)
    from emsim_amd.backend.slice_builder_hip import OneSliceBuilder
    b = OneSliceBuilder([6, 79], 128, 128, 0.5, 0.5)
    got = host(b.bin_atoms_one_slice(pts, counts))
    ref = orc.bin_atoms_one_slice(pts, counts, 128, 128, 0.5, 0.5)
    np.testing.assert_array_equal(got, ref)


def test_binning_is_deterministic_and_empty_input_ok():
    from emsim_amd.backend.slice_builder_hip import MultiSlicesBuilder
    b = MultiSlicesBuilder([6], 3, 32, 32, 1.0, 1.0, 1.0)
    h = host(b.bin_atoms_multi_slices(np.zeros((0, 3), np.float32), [0]))
    assert h.shape == (1, 3, 32, 32) and not h.any()
    pts = _edge_cloud((3, 32, 32), (1.0, 1.0, 1.0), 3000, 5)
    a = host(b.bin_atoms_multi_slices(pts, [3000]))
    c = host(b.bin_atoms_multi_slices(pts, [3000]))
    np.testing.assert_array_equal(a, c)


# ---------------------------------------------------------------------------
# potential slices
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,ns,bandlimit", [(64, 5, True), (64, 3, False), (256, 45, True),
                                            (512, 4, True), (16, 2, True)])
def test_make_multi_slices(n, ns, bandlimit):
    rng = np.random.default_rng(n + ns)
    zs = [1, 6, 7, 8, 16]
    hist = (rng.random((5, ns, n, n)) < 0.01).astype(np.float32) * rng.integers(1, 4, (5, ns, n, n))
    hist = hist.astype(np.float32)
    sf = orc.scattering_factors_half(zs, n, n, 1.0, 1.0)
    from emsim_amd.backend.slice_builder_hip import MultiSlicesBuilder
    b = MultiSlicesBuilder(zs, ns, n, n, 2.0, 1.0, 1.0)
    np.testing.assert_array_equal(b.scattering_factors, sf)
    got = host(b.make_multi_slices(hist, symmetric_bandlimit=bandlimit))
    ref32 = orc.make_multi_slices_fftw32(hist, sf, bandlimit)
    ref64 = orc.make_multi_slices_f64(hist, sf, bandlimit)
    e32, e64 = report(f"slices n={n} ns={ns} bl={bandlimit}", got, ref32, ref64)
    assert e32 < TOL and e64 < TOL64
    assert got.min() >= np.float32(1e-13)


def test_make_slices_rectangular_and_bandlimit_rule():
    rng = np.random.default_rng(3)
    zs = [6, 8]
    hist = (rng.random((2, 3, 32, 64)) < 0.02).astype(np.float32)
    from emsim_amd.backend.slice_builder_hip import MultiSlicesBuilder
    b = MultiSlicesBuilder(zs, 3, 32, 64, 1.0, 0.5, 0.25)
    sf = orc.scattering_factors_half(zs, 32, 64, 0.5, 0.25)
    got = host(b.make_multi_slices(hist, symmetric_bandlimit=False))
    ref32 = orc.make_multi_slices_fftw32(hist, sf, False)
    assert orc.rel_l2(got, ref32) < TOL
    with pytest.raises(ValueError):
        b.make_multi_slices(hist, symmetric_bandlimit=True)


def test_build_multi_slices_protein_c2():
    mol = atm.centralize(synth.protein_cloud(50000, 0))
    got = host(pot.build_multi_slices(mol, dz=2.0, pixel_size=1.0, lateral_size=256))
    ref32, ref64, _ = slices_oracle(mol, 2.0, 1.0, 256)
    assert got.shape == ref32.shape == (45, 256, 256)
    e32, e64 = report("build_multi_slices C2", got, ref32, ref64)
    assert e32 < TOL and e64 < TOL64


def test_build_one_slice_five_atoms():
    mol = atm.centralize(synth.five_atoms())
    pixel = 50 / 512
    got = host(pot.build_one_slice(mol, pixel_size=pixel, lateral_size=(512, 512)))
    m, zs, counts = atm.sort_elements_and_count(mol)
    hist = orc.bin_atoms_one_slice(m.coordinates, counts, 512, 512, pixel, pixel)
    sf = orc.scattering_factors_half(list(zs), 512, 512, pixel, pixel)
    ref32 = orc.make_one_slice_fftw32(hist, sf)
    ref64 = orc.make_one_slice_f64(hist, sf)
    e32, e64 = report("build_one_slice C1", got, ref32, ref64)
    assert e32 < TOL and e64 < TOL64


# ---------------------------------------------------------------------------
# wave propagation
# ---------------------------------------------------------------------------
def _prop(n, d=1.0):
    return emsim_amd.wave.get_raw_wave_propagator(n, d, 200)


def _wave(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)


@pytest.mark.parametrize("n", [8, 64, 256, 512, 1024])
def test_space_propagate(n):
    p = _prop(n)
    w = _wave(n, n)
    got = host(p.space_propagate(w, 2.0))
    ref32 = orc.space_propagate_fftw32(w, 2.0, LAM, 1.0, 1.0)
    ref64 = orc.space_propagate_f64(w, 2.0, LAM, 1.0, 1.0)
    e32, e64 = report(f"space_propagate n={n}", got, ref32, ref64)
    assert e32 < TOL and e64 < TOL64


def test_slice_transmit_and_init_wave():
    n = 128
    p = _prop(n, 0.5)
    w0 = host(p.init_wave(20.0))
    np.testing.assert_array_equal(w0, orc.init_wave(n, n, 0.5, 0.5, 20.0))
    v = (np.random.default_rng(1).random((n, n)) * 30).astype(np.float32)
    w = _wave(n, 2)
    got = host(p.slice_transmit(w, v))
    ref32 = orc.slice_transmit_fftw32(w, v, LAM, GAMMA)
    ref64 = orc.slice_transmit_f64(w, v, LAM, GAMMA)
    e32, e64 = report("slice_transmit", got, ref32, ref64)
    assert e32 < 1e-6 and e64 < 1e-6


@pytest.mark.parametrize("n,d,cs,df,ap", [(512, 50 / 512, 1.3, 699.32, 10.37e-3),
                                          (256, 1.0, 1.3, 700.0, 10e-3),
                                          (64, 1.0, 0.0, 0.0, 1.0)])
def test_lens_propagate_and_intensity(n, d, cs, df, ap):
    p = _prop(n, d)
    w = _wave(n, 7)
    got = host(p.lens_propagate(w, cs, df, ap))
    ref32 = orc.lens_propagate_fftw32(w, LAM, cs, df, ap, d, d)
    ref64 = orc.lens_propagate_f64(w, LAM, cs, df, ap, d, d)
    e32, e64 = report(f"lens n={n}", got, ref32, ref64)
    assert e32 < TOL and e64 < TOL64
    img = host(p.lens_intensity(w, cs, df, ap))
    assert orc.rel_l2(img, orc.intensity(ref32)) < TOL


@pytest.mark.parametrize("n,ns", [(256, 45), (512, 20), (128, 3)])
def test_multislice_propagate(n, ns):
    rng = np.random.default_rng(ns)
    v = (rng.random((ns, n, n)) * 15).astype(np.float32)
    p = _prop(n)
    w0 = orc.init_wave(n, n, 1.0, 1.0, 20.0)
    got = host(p.multislice_propagate(w0, v, 2.0))
    ref32 = orc.multislice_fftw32(w0, v, 2.0, LAM, GAMMA, 1.0, 1.0)
    ref64 = orc.multislice_f64(w0, v, 2.0, LAM, GAMMA, 1.0, 1.0)
    e32, e64 = report(f"multislice n={n} ns={ns}", got, ref32, ref64)
    assert e32 < TOL and e64 < TOL64
    one = host(p.singleslice_propagate(w0, v[0], 2.0))
    ref1 = orc.space_propagate_fftw32(orc.slice_transmit_fftw32(w0, v[0], LAM, GAMMA), 2.0,
                                      LAM, 1.0, 1.0)
    assert orc.rel_l2(one, ref1) < TOL


def test_batched_calls_equal_single_calls():
    n, ns, b = 128, 6, 5
    rng = np.random.default_rng(9)
    v = (rng.random((b, ns, n, n)) * 15).astype(np.float32)
    p = _prop(n)
    w0 = p.init_wave(20.0, batch=b)
    out = host(p.multislice_propagate(w0, v, 2.0))
    for i in range(b):
        single = host(p.multislice_propagate(w0[i], v[i], 2.0))
        np.testing.assert_array_equal(out[i], single)
    img = host(p.lens_intensity(out, 1.3, 700.0, 10e-3))
    for i in range(b):
        np.testing.assert_array_equal(img[i], host(p.lens_intensity(out[i], 1.3, 700.0, 10e-3)))


def test_bad_arguments_raise():
    p = _prop(64)
    with pytest.raises(ValueError):
        p.space_propagate(np.zeros((32, 32), np.complex64), 1.0)
    with pytest.raises(TypeError):
        p.space_propagate(np.zeros((64, 64), np.float32), 1.0)
    # sizes outside 8..4096 (powers of two) and 8..2048 (chirp-z)
    for n in (4, 3000, 8192):
        with pytest.raises(ValueError):
            emsim_amd.wave.get_raw_wave_propagator(n, 1.0, 200)
    from emsim_amd.backend.slice_builder_hip import MultiSlicesBuilder
    with pytest.raises(ValueError):
        MultiSlicesBuilder([6], 3, 3000, 3000, 1.0, 1.0, 1.0)
    # numpy.arange makes 9 edges for 7 slices of 0.3 A: the reference would
    # build an 8-slice histogram and misalign it; the hip builder refuses
    with pytest.raises(ValueError, match="edges"):
        MultiSlicesBuilder([6], 7, 64, 64, 0.3, 1.0, 1.0)


# ---------------------------------------------------------------------------
# end to end (the reference's demos / Pipe)
# ---------------------------------------------------------------------------
def test_five_atoms_demo_c1():
    """demos/five_atoms.py: build_one_slice -> slice_transmit -> lens -> |psi|^2."""
    tem = emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                            aperture=10.37e-3)
    tem.set_scherzer_condition()
    mol = atm.centralize(synth.five_atoms())
    shape, pixel = (512, 512), 50 / 512
    prop = tem.get_wave_propagator(wave_shape=shape, pixel_size=pixel)
    aslice = pot.build_one_slice(mol, pixel_size=pixel, lateral_size=shape)
    exit_wave = prop.slice_transmit(prop.init_wave(), aslice)
    image = host(prop.raw_propagator.lens_intensity(exit_wave, tem.cs_mm, tem.defocus,
                                                    tem.aperture))
    # oracle chain
    m, zs, counts = atm.sort_elements_and_count(mol)
    hist = orc.bin_atoms_one_slice(m.coordinates, counts, 512, 512, pixel, pixel)
    sf = orc.scattering_factors_half(list(zs), 512, 512, pixel, pixel)
    v32, v64 = orc.make_one_slice_fftw32(hist, sf), orc.make_one_slice_f64(hist, sf)
    w0 = orc.init_wave(512, 512, pixel, pixel, 20)
    i32 = orc.intensity(orc.lens_propagate_fftw32(
        orc.slice_transmit_fftw32(w0, v32.astype(np.float32), LAM, GAMMA),
        LAM, tem.cs_mm, tem.defocus, tem.aperture, pixel, pixel))
    i64 = np.abs(orc.lens_propagate_f64(orc.slice_transmit_f64(w0, v64, LAM, GAMMA), LAM,
                                        tem.cs_mm, tem.defocus, tem.aperture, pixel, pixel)) ** 2
    e32, e64 = report("five atoms image", image, i32, i64)
    assert e32 < TOL and e64 < TOL64


def test_pipe_run_protein_c2():
    tem = emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                            aperture=10e-3)
    pipe = emsim_amd.pipe.Pipe(tem, resolution=2.0, slice_thickness=2.0, roi=256,
                               add_water=False)
    mol = synth.protein_cloud(50000, 0)
    image = host(pipe.run(mol))
    exit_wave = host(pipe.get_exit_wave(mol))
    i32, i64, x32, x64 = pipe_oracle(mol, tem, 2.0, 2.0, (256, 256))
    e32, e64 = report("Pipe.run C2 exit wave", exit_wave, x32, x64)
    assert e32 < TOL and e64 < TOL64
    e32, e64 = report("Pipe.run C2 image", image, i32, i64)
    assert e32 < TOL and e64 < TOL64


def test_add_water_statistics_and_vacancy_rule():
    from emsim_amd.backend.slice_builder_hip import MultiSlicesBuilder
    ns, n = 8, 128
    b = MultiSlicesBuilder([1, 6, 8], ns, n, n, 2.0, 1.0, 1.0)
    hist = np.zeros((3, ns, n, n), np.float32)
    hist[1, :, :64] = 1.0                      # carbon fills half of every slice
    out = host(b.add_water(hist, seed=1234))
    lam = emsim_amd.physics.water_num_dens * 1.0 * 1.0 * 2.0
    occupied = hist.sum(0) > 0
    assert not out[0][occupied].any() and not out[2][occupied].any()
    np.testing.assert_array_equal(out[1], hist[1])
    o, h = out[2][~occupied], out[0][~occupied]
    nvox = o.size
    assert abs(o.mean() - lam) < 5 * np.sqrt(lam / nvox)
    assert abs(h.mean() - 2 * lam) < 5 * np.sqrt(2 * lam / nvox)
    assert abs(o.var() - lam) < 0.05 * lam
    again = host(b.add_water(hist, seed=1234))
    np.testing.assert_array_equal(out, again)
