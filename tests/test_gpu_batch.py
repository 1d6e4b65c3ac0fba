"""GPU parity of the fused batched path (emsim_amd.batch.BatchPipe).

Each image must match the oracle chain of the reference's per-image path
(pot.build_multi_slices(..., n_slices) -> init_wave -> multislice_propagate
-> lens_propagate -> |psi|^2, fftw semantics) within 1e-5 relative L2, and
the fused kernels must agree with the reference-shaped kernels.
"""
import numpy as np
import pytest
import torch

import emsim_amd
from emsim_amd import atoms as atm
from emsim_amd import synth
from emsim_amd.batch import BatchPipe
from emsim_amd.utils.rot import random_uniform_quaternions
from oracle import emsim_oracle as orc
from tests._chain import slices_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5
ELEMS = [1, 6, 7, 8, 16]


def _tem():
    return emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                             aperture=10e-3)


def _mols(sizes, seed):
    quats = random_uniform_quaternions(len(sizes), rng=np.random.default_rng(seed))
    return [synth.rotated(synth.protein_cloud(int(n), seed + i), quats[i])
            for i, n in enumerate(sizes)]


def _oracle_image(mol, tem, n, ns, dz, pixel):
    mol = atm.centralize(mol)
    v32, v64, _ = slices_oracle(mol, dz, pixel, (n, n), ns)
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    w0 = orc.init_wave(n, n, pixel, pixel, tem.electron_dose)
    x32 = orc.multislice_fftw32(w0, v32.astype(np.float32), dz, lam, gamma, pixel, pixel)
    x64 = orc.multislice_f64(w0, v64, dz, lam, gamma, pixel, pixel)
    i32 = orc.intensity(orc.lens_propagate_fftw32(x32, lam, tem.cs_mm, tem.defocus,
                                                  tem.aperture, pixel, pixel))
    i64 = np.abs(orc.lens_propagate_f64(x64, lam, tem.cs_mm, tem.defocus, tem.aperture,
                                        pixel, pixel)) ** 2
    return i32, i64, x32, x64


@pytest.fixture(autouse=True)
def _gpu(hip):
    yield
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,ns,dz,pixel", [(256, 24, 2.0, 1.0), (128, 10, 1.0, 0.5)])
def test_fused_batch_matches_oracle(n, ns, dz, pixel):
    tem = _tem()
    mols = _mols([4000, 9000, 15000], 31)
    pipe = BatchPipe(tem, pixel, dz, n, ns, ELEMS)
    batch = pipe.upload([pipe.prepare(m) for m in mols])
    images = pipe.run(batch).cpu().numpy()
    waves = pipe.exit_waves(batch).cpu().numpy()
    for b, mol in enumerate(mols):
        i32, i64, x32, x64 = _oracle_image(mol, tem, n, ns, dz, pixel)
        ei, ei64 = orc.rel_l2(images[b], i32), orc.rel_l2(images[b], i64)
        ex, ex64 = orc.rel_l2(waves[b], x32), orc.rel_l2(waves[b], x64)
        print(f"[fused n={n} ns={ns} mol {b}] image {ei:.2e} (fp64 {ei64:.2e}), "
              f"exit wave {ex:.2e} (fp64 {ex64:.2e})", flush=True)
        assert ei < TOL and ei64 < TOL and ex < TOL and ex64 < TOL


def _sparse_planes_molecule(seed=7):
    """Element planes of every kind the spectrum pass distinguishes, at
    n = 128, dz = 1, 12 slices: empty slices (0, 1 and 11), an element
    absent throughout (O), planes of 1, 2, 8 atoms (direct 2-D DFT) and of 9
    and tens to hundreds of atoms (row DFTs + column transform).  Coordinates are
    (z, x, y); the bounding box is centred by two H atoms at its corners."""
    rng = np.random.default_rng(seed)

    def cloud(n, z0, z1):
        z = rng.uniform(z0 + 0.05, z1 - 0.05, n)
        xy = rng.uniform(-39.0, 39.0, (n, 2))
        return np.column_stack([z, xy])

    parts = [(1, np.array([[-4.0, -40.0, -40.0], [4.0, 40.0, 40.0]])),
             (1, cloud(300, -4.0, 4.0)),
             (6, cloud(600, -4.0, -1.0)),
             (7, cloud(8, 0.0, 1.0)), (7, cloud(9, 1.0, 2.0)),
             (16, cloud(1, -1.0, 0.0)), (16, cloud(2, 2.0, 3.0))]
    el = np.concatenate([np.full(len(c), z, np.int64) for z, c in parts])
    xyz = np.concatenate([c for _, c in parts])
    return atm.AtomList(elements=el, coordinates=xyz, label="sparse planes")


@pytest.mark.parametrize("n,ns,dz,pixel", [(64, 3, 12.0, 1.0), (1024, 3, 12.0, 0.5),
                                           (4096, 1, 36.0, 0.25)])
def test_fused_other_sizes_match_oracle(n, ns, dz, pixel):
    """Sizes whose spectrum pass takes other code: 1024 (row table beyond
    the LDS stage: bucket pieces), 4096 (large pieces, column n2/2 folded
    into column 0, 16 values per thread), 64 (one block holds several
    tiles' worth of rows)."""
    tem = _tem()
    mol = _mols([3000], 81)[0]
    pipe = BatchPipe(tem, pixel, dz, n, ns, ELEMS)
    batch = pipe.upload([pipe.prepare(mol)])
    image = pipe.run(batch).cpu().numpy()[0]
    wave = pipe.exit_waves(batch).cpu().numpy()[0]
    i32, i64, x32, x64 = _oracle_image(mol, tem, n, ns, dz, pixel)
    ei, ex = orc.rel_l2(image, i32), orc.rel_l2(wave, x32)
    ei64, ex64 = orc.rel_l2(image, i64), orc.rel_l2(wave, x64)
    print(f"[fused n={n}] image {ei:.2e} (fp64 {ei64:.2e}), exit wave {ex:.2e} "
          f"(fp64 {ex64:.2e})", flush=True)
    assert ei < TOL and ei64 < TOL and ex < TOL and ex64 < TOL


def test_sparse_element_planes_match_oracle():
    tem = _tem()
    n, ns, dz, pixel = 128, 12, 1.0, 1.0
    mol = _sparse_planes_molecule()
    pipe = BatchPipe(tem, pixel, dz, n, ns, ELEMS)
    batch = pipe.upload([pipe.prepare(mol)])
    images = pipe.run(batch).cpu().numpy()
    waves, slices = pipe.exit_waves_and_slices(batch)
    i32, i64, x32, x64 = _oracle_image(mol, tem, n, ns, dz, pixel)
    ei, ex = orc.rel_l2(images[0], i32), orc.rel_l2(waves[0].cpu().numpy(), x32)
    ei64, ex64 = orc.rel_l2(images[0], i64), orc.rel_l2(waves[0].cpu().numpy(), x64)
    ref = pipe.build_slices(batch).cpu().numpy()
    got = slices.cpu().numpy()
    es = orc.rel_l2(got, ref)
    print(f"[sparse planes] image {ei:.2e} (fp64 {ei64:.2e}), exit wave {ex:.2e} "
          f"(fp64 {ex64:.2e}), slices vs builder {es:.2e}", flush=True)
    assert ei < TOL and ei64 < TOL and ex < TOL and ex64 < TOL and es < 2e-6
    # empty slices are the clip value exactly; each slice with atoms matches too
    for s in (0, 1, 11):
        assert np.all(got[0, s] == np.float32(1e-13))
    for s in range(2, 11):
        assert orc.rel_l2(got[0, s], ref[0, s]) < 2e-6


def test_dense_slices_match_oracle():
    """Slices with more atoms than the spectrum kernel stages at once
    (> 8192 per molecule and slice): the row DFTs take the bucket in pieces.
    A 256^2 random slab, 40k atoms in 4 slices of 2 A (about 10k each)."""
    tem = _tem()
    rng = np.random.default_rng(5)
    n_atoms = 40000
    zs = rng.choice([1, 6, 8], size=n_atoms, p=[0.6, 0.2, 0.2])
    xyz = np.column_stack([rng.uniform(-3.95, 3.95, n_atoms),
                           rng.uniform(-120.0, 120.0, (n_atoms, 2))])
    mol = atm.AtomList(elements=zs, coordinates=xyz, label="dense slab")
    n, ns, dz, pixel = 256, 4, 2.0, 1.0
    pipe = BatchPipe(tem, pixel, dz, n, ns, ELEMS)
    batch = pipe.upload([pipe.prepare(mol)])
    images = pipe.run(batch).cpu().numpy()
    waves, slices = pipe.exit_waves_and_slices(batch)
    i32, i64, x32, x64 = _oracle_image(mol, tem, n, ns, dz, pixel)
    ei, ex = orc.rel_l2(images[0], i32), orc.rel_l2(waves[0].cpu().numpy(), x32)
    es = orc.rel_l2(slices.cpu().numpy(), pipe.build_slices(batch).cpu().numpy())
    print(f"[dense slices] image {ei:.2e}, exit wave {ex:.2e}, slices vs builder {es:.2e}",
          flush=True)
    assert ei < TOL and ex < TOL and es < 2e-6


def test_bench_workload_images_match_port():
    """The benchmark's own configuration -- 512^2, 100 slices of 1 A, its batch
    of 64 molecules (per-slice spectrum order, the propagator's dead columns
    left out between passes) -- against oracle/fftw_port.c for the largest
    and the smallest molecule of the batch."""
    port = pytest.importorskip("oracle.port")
    try:
        port.load()
    except Exception as exc:
        pytest.fail(f"oracle/fftw_port is not built: {exc}")
    import bench
    tem = emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                            aperture=10e-3)
    elements, n, ns, d, dz = [1, 6, 7, 8, 16], 512, 100, 1.0, 1.0
    mols, _ = bench.workload(64)
    pipe = BatchPipe(tem, d, dz, n, ns, elements)
    images = pipe.run(pipe.upload([pipe.prepare(m) for m in mols])).cpu().numpy()
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    psi0 = float(np.float32(np.sqrt(tem.electron_dose * d * d)))
    sf = orc.scattering_factors_half(elements, n, n, d, d)
    sizes = [len(m) for m in mols]
    for i in (int(np.argmax(sizes)), int(np.argmin(sizes))):
        mol, _, counts = atm.sort_elements_and_count(atm.centralize(mols[i]), elements)
        ref = port.simulate(mol.coordinates, counts, ns, n, dz, d, sf, lam, gamma, tem.cs_mm,
                            tem.defocus, tem.aperture, psi0)
        e = orc.rel_l2(images[i], ref)
        print(f"bench molecule {i} ({sizes[i]} atoms): image rel L2 vs fftw_port {e:.2e}")
        assert e < TOL


def test_batch_sim_delivers_in_order_and_skips_failures():
    """The batched simulation loop: batches of 3 over 7 molecules (the last
    batch partial), images through the pinned ring equal to one fused run of
    all molecules, input order and labels kept, a molecule with an element
    outside the pipe's list logged and skipped."""
    from emsim_amd.batch import BatchSim
    tem = _tem()
    mols = _mols([2000, 3000, 2500, 4000, 1500, 3500, 2200], 71)
    for i, m in enumerate(mols):
        m.label = f"mol{i}"
    bad = atm.AtomList(elements=np.array([79, 6]), coordinates=np.zeros((2, 3)), label="gold")
    pipe = BatchPipe(tem, 1.0, 2.0, 128, 8, ELEMS)
    ref = pipe.run(pipe.upload([pipe.prepare(m) for m in mols])).cpu().numpy()
    got = []
    sim = BatchSim(pipe, mols[:3] + [bad] + mols[3:], lambda img, label: got.append((label, img)),
                   batch=3, depth=2)
    assert sim.run() == 7 and sim.failed == 1
    assert [label for label, _ in got] == [f"mol{i}" for i in range(7)]
    for i, (_, img) in enumerate(got):
        np.testing.assert_array_equal(img, ref[i])


def test_fused_equals_unfused_and_is_deterministic():
    tem = _tem()
    mols = _mols([3000, 6000], 41)
    pipe = BatchPipe(tem, 1.0, 2.0, 128, 12, ELEMS)
    batch = pipe.upload([pipe.prepare(m) for m in mols])
    a = pipe.run(batch)
    b = pipe.run(batch)
    assert torch.equal(a, b)
    u = pipe.run_unfused(batch)
    assert orc.rel_l2(a.cpu().numpy(), u.cpu().numpy()) < 2e-6


def test_batched_equals_single_and_empty_molecule():
    tem = _tem()
    mols = _mols([3000, 5000], 51)
    pipe = BatchPipe(tem, 1.0, 2.0, 128, 8, ELEMS)
    prepared = [pipe.prepare(m) for m in mols]
    empty = emsim_amd.batch.PreparedMolecule(np.zeros((0, 3), np.float32),
                                             np.zeros(5, np.int64), "empty")
    together = pipe.run(pipe.upload([prepared[0], empty, prepared[1]])).cpu().numpy()
    alone0 = pipe.run(pipe.upload([prepared[0]])).cpu().numpy()
    alone1 = pipe.run(pipe.upload([prepared[1]])).cpu().numpy()
    np.testing.assert_array_equal(together[0], alone0[0])
    np.testing.assert_array_equal(together[2], alone1[0])
    # no atoms: V = 1e-13 everywhere (the clip), so the image is the lens
    # acting on an (essentially) plane wave: uniform dose d1 d2
    assert abs(together[1].mean() - 20.0) / 20.0 < 1e-4


def test_fused_slices_output_matches_slice_builder():
    """The potentials the fused path can write on the way equal the
    reference-shaped builder's, and writing them leaves the waves unchanged."""
    tem = _tem()
    mols = _mols([4000, 7000], 61)
    pipe = BatchPipe(tem, 1.0, 8.0, 128, 6, ELEMS)
    batch = pipe.upload([pipe.prepare(m) for m in mols])
    waves, slices = pipe.exit_waves_and_slices(batch)
    ref = pipe.build_slices(batch)
    err = orc.rel_l2(slices.cpu().numpy(), ref.cpu().numpy())
    print(f"fused slices vs builder: {err:.2e}")
    assert err < 2e-6
    assert torch.equal(waves, pipe.exit_waves(batch))


def test_graph_replay_equals_eager_and_follows_new_coords():
    """A captured HIP graph of the fused path replays to the eager result
    (several replays, eager work in between) and follows atoms overwritten
    in place."""
    tem = _tem()
    mols = _mols([3000, 6000, 4500], 71)
    pipe = BatchPipe(tem, 1.0, 2.0, 128, 10, ELEMS)
    prepared = [pipe.prepare(m) for m in mols]
    batch = pipe.upload(prepared)
    eager = pipe.run(batch).clone()
    g = pipe.capture(batch)
    for _ in range(3):
        out = g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
        pipe.run(batch)
        torch.cuda.synchronize()
    quats = random_uniform_quaternions(len(mols), rng=np.random.default_rng(72))
    prepared2 = [pipe.prepare(synth.rotated(m, q)) for m, q in zip(mols, quats)]
    assert all((a.counts == b.counts).all() for a, b in zip(prepared, prepared2))
    batch2 = pipe.upload(prepared2)
    g.coords.copy_(batch2.coords)
    out2 = g.replay().clone()
    torch.cuda.synchronize()
    assert torch.equal(out2, pipe.run(batch2))
    assert not torch.equal(out2, eager)


@pytest.mark.parametrize("n", [128, 256])
def test_launch_shapes_do_not_change_results(n):
    """Every launch-shape setting, column-tile shapes included (256 has the
    alternative shapes), gives bit-identical images."""
    from emsim_amd import autotune
    tem = _tem()
    mols = _mols([3000, 5000], 81)
    pipe = BatchPipe(tem, 1.0, 2.0, n, 8, ELEMS)
    batch = pipe.upload([pipe.prepare(m) for m in mols])
    autotune.apply({})
    ref = pipe.run(batch).clone()
    try:
        for cfg in ({"row_items_per_block": 1, "col_tiles_per_block": 1, "col_shape": 1},
                    {"row_items_per_block": 7, "col_tiles_per_block": 5, "col_shape": 2},
                    {"row_items_per_block": 16, "col_tiles_per_block": 8, "col_shape": 0}):
            autotune.apply(cfg)
            assert autotune.current() == cfg
            assert torch.equal(pipe.run(batch), ref)
        best, table = autotune.tune(lambda: pipe.run(batch), reps=1)
        print("tuned", best, "from", len(table), "settings")
        assert torch.equal(pipe.run(batch), ref)
    finally:
        autotune.apply({})


def test_batched_wpo_equals_pipe_run_wpo():
    """Batched weak-phase-object images against the reference-API path
    (Pipe.run_wpo: build_one_slice -> singleslice_propagate(dz=0) -> lens)
    molecule by molecule: the reference-shaped batch (same kernels) bit for
    bit, the fused batch (potential from the atom buckets) within 2e-6."""
    emsim_amd.config.set_backend("hip")
    tem = _tem()
    mols = _mols([2500, 6000, 4000], 91)
    pipe = BatchPipe(tem, 1.0, 2.0, 128, 8, ELEMS)
    batch = pipe.upload([pipe.prepare(m) for m in mols])
    fused = pipe.run_wpo(batch)
    unfused = pipe.run_wpo_unfused(batch)
    ref_pipe = emsim_amd.pipe.Pipe(tem, resolution=2.0, slice_thickness=2.0, roi=128,
                                   add_water=False)
    for b, m in enumerate(mols):
        ref = ref_pipe.run_wpo(m)
        ref = ref if isinstance(ref, torch.Tensor) else torch.as_tensor(ref)
        ref = ref.to(unfused.device).reshape(unfused[b].shape)
        assert torch.equal(unfused[b], ref)
        assert orc.rel_l2(fused[b].cpu().numpy(), ref.cpu().numpy()) < 2e-6


def test_orientations_on_device():
    """One molecule in several orientations rotated and centred on the GPU:
    coordinates agree with the host rotation to float32 rounding, the images
    with the host-prepared batch, and each image with the oracle chain run
    on the device-made coordinates."""
    tem = _tem()
    mol = synth.protein_cloud(5000, 101)
    quats = random_uniform_quaternions(4, rng=np.random.default_rng(5))
    n, ns, dz, pixel = 128, 10, 2.0, 1.0
    pipe = BatchPipe(tem, pixel, dz, n, ns, ELEMS)
    dev_batch = pipe.upload_orientations(mol, quats)
    host = pipe.upload([pipe.prepare(synth.rotated(mol, q)) for q in quats])
    assert np.array_equal(dev_batch.mol_off, host.mol_off)
    assert np.array_equal(dev_batch.counts, host.counts)
    diff = (dev_batch.coords - host.coords).abs().max().item()
    print(f"device vs host rotation: max coordinate difference {diff:.2e} A")
    assert diff < 1e-3
    img_dev = pipe.run(dev_batch).cpu().numpy()
    img_host = pipe.run(host).cpu().numpy()
    assert orc.rel_l2(img_dev, img_host) < 1e-3
    coords = dev_batch.coords.cpu().numpy()
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    for b in (0, 3):
        c = coords[dev_batch.mol_off[b]:dev_batch.mol_off[b + 1]]
        z = np.repeat(np.asarray(ELEMS), dev_batch.counts[b])
        v32, _, _ = slices_oracle(atm.AtomList(z, c), dz, pixel, (n, n), ns)
        w0 = orc.init_wave(n, n, pixel, pixel, tem.electron_dose)
        x32 = orc.multislice_fftw32(w0, v32.astype(np.float32), dz, lam, gamma, pixel, pixel)
        i32 = orc.intensity(orc.lens_propagate_fftw32(x32, lam, tem.cs_mm, tem.defocus,
                                                      tem.aperture, pixel, pixel))
        err = orc.rel_l2(img_dev[b], i32)
        print(f"orientation {b}: image vs oracle {err:.2e}")
        assert err < TOL
