"""GPU parity of the weak-phase-object path, the threaded EMSim loop and the
mixed-size batch at its full workload.

Weak-phase-object images (emsim/pipe.py:57-64,73-77): build_one_slice ->
init_wave -> singleslice_propagate(dz = 0) -> lens -> |psi|^2, through the
reference API (Pipe.run_wpo) and the batched fused path (BatchPipe.run_wpo),
against the oracle chain at the five-atom (512^2) and protein-cloud (256^2)
sizes.

EMSim (emsim/simulator.py:21-91) drives the hip Pipe from its consumer thread
on a real HIP stream; the numpy images it delivers are checked against the
oracle chain, and a molecule that fails is reported and skipped.

Mixed sizes (256 molecules of 5k-200k atoms, 128^2 / 256^2 / 512^2 waves,
1 A pixels, 2 A slices): here 64 molecules drawn the same way; one molecule
of every wave size, and the largest molecule on a 128^2 wave (where atoms
fall outside the window), against oracle/fftw_port.c; images in the mixed
batch equal the same molecules simulated alone.
"""
import numpy as np
import pytest
import torch

import emsim_amd
from emsim_amd import atoms as atm
from emsim_amd import synth
from emsim_amd.batch import BatchPipe, MixedBatch, pool_map
from emsim_amd.utils.rot import random_uniform_quaternions
from oracle import emsim_oracle as orc
from tests._chain import pipe_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5
ELEMS = [1, 6, 7, 8, 16]


@pytest.fixture(autouse=True)
def _gpu(hip):
    emsim_amd.config.set_backend("hip")
    yield
    torch.cuda.synchronize()


def _tem(aperture=10e-3, defocus=700.0):
    return emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3,
                             defocus=defocus, aperture=aperture)


def wpo_oracle(mol, tem, pixel, n, elements=None):
    """(image fftw32, image f64) of Pipe.run_wpo: one-slice potential,
    transmission, dz = 0 propagation (the 2/3 mask alone), lens, |psi|^2."""
    mol = atm.centralize(mol)
    m, zs, counts = atm.sort_elements_and_count(mol, elements or [])
    hist = orc.bin_atoms_one_slice(m.coordinates, counts, n, n, pixel, pixel)
    sf = orc.scattering_factors_half(list(zs), n, n, pixel, pixel)
    v32, v64 = orc.make_one_slice_fftw32(hist, sf), orc.make_one_slice_f64(hist, sf)
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    w0 = orc.init_wave(n, n, pixel, pixel, tem.electron_dose)
    x32 = orc.space_propagate_fftw32(orc.slice_transmit_fftw32(w0, v32.astype(np.float32), lam,
                                                               gamma), 0.0, lam, pixel, pixel)
    x64 = orc.space_propagate_f64(orc.slice_transmit_f64(w0, v64, lam, gamma), 0.0, lam, pixel,
                                  pixel)
    i32 = orc.intensity(orc.lens_propagate_fftw32(x32, lam, tem.cs_mm, tem.defocus, tem.aperture,
                                                  pixel, pixel))
    i64 = np.abs(orc.lens_propagate_f64(x64, lam, tem.cs_mm, tem.defocus, tem.aperture, pixel,
                                        pixel)) ** 2
    return i32, i64


def _host(x):
    return x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def test_wpo_five_atoms_512_matches_oracle():
    """C1 geometry and Scherzer optics: Pipe.run_wpo and the batched fused
    path against the oracle chain."""
    tem = _tem()
    tem.set_scherzer_condition()
    mol = synth.five_atoms()
    pixel = 50 / 512
    i32, i64 = wpo_oracle(mol, tem, pixel, 512)
    pipe = emsim_amd.pipe.Pipe(tem, resolution=2 * pixel, slice_thickness=2.0, roi=512,
                               add_water=False)
    got = _host(pipe.run_wpo(mol))
    e32, e64 = orc.rel_l2(got, i32), orc.rel_l2(got, i64)
    print(f"[wpo C1] Pipe.run_wpo vs fftw32 {e32:.2e}, vs fp64 {e64:.2e}", flush=True)
    assert e32 < TOL and e64 < TOL
    elems = [6, 14, 29, 79, 92]
    bp = BatchPipe(tem, pixel, 2.0, 512, 1, elems)
    batch = bp.upload([bp.prepare(mol)])
    fused = _host(bp.run_wpo(batch)[0])
    ef = orc.rel_l2(fused, i32)
    print(f"[wpo C1] BatchPipe.run_wpo vs fftw32 {ef:.2e}", flush=True)
    assert ef < TOL
    unf = _host(bp.run_wpo_unfused(batch)[0])
    assert orc.rel_l2(fused, unf) < 2e-6


def test_wpo_protein_256_batch_matches_oracle():
    """C2-style clouds at 256^2 (1 A pixels): three molecules in one batched
    fused launch sequence, each against the oracle chain and against
    Pipe.run_wpo; an empty molecule in the batch images a plane wave."""
    tem = _tem()
    quats = random_uniform_quaternions(3, rng=np.random.default_rng(301))
    mols = [synth.rotated(synth.protein_cloud(n, 300 + i), quats[i])
            for i, n in enumerate((50000, 12000, 30000))]
    bp = BatchPipe(tem, 1.0, 2.0, 256, 1, ELEMS)
    prepared = [bp.prepare(m) for m in mols]
    empty = emsim_amd.batch.PreparedMolecule(np.zeros((0, 3), np.float32),
                                             np.zeros(len(ELEMS), np.int64), "empty")
    images = _host(bp.run_wpo(bp.upload(prepared[:1] + [empty] + prepared[1:])))
    pipe = emsim_amd.pipe.Pipe(tem, resolution=2.0, slice_thickness=2.0, roi=256,
                               add_water=False)
    for b, mol in zip((0, 2, 3), mols):
        i32, i64 = wpo_oracle(mol, tem, 1.0, 256, ELEMS)
        ef, e64 = orc.rel_l2(images[b], i32), orc.rel_l2(images[b], i64)
        ep = orc.rel_l2(_host(pipe.run_wpo(mol)), i32)
        print(f"[wpo 256 mol {b}] fused vs fftw32 {ef:.2e} (fp64 {e64:.2e}), "
              f"Pipe.run_wpo vs fftw32 {ep:.2e}", flush=True)
        assert ef < TOL and e64 < TOL and ep < TOL
    assert abs(images[1].mean() - 20.0) / 20.0 < 1e-4


def test_emsim_drives_hip_pipe_on_consumer_thread():
    """The reference's threaded loop with the hip Pipe: images arrive as numpy
    arrays in input order, equal to the oracle chain; a molecule that fails
    (an element without form-factor parameters) is logged and skipped."""
    tem = _tem()
    pipe = emsim_amd.pipe.Pipe(tem, resolution=2.0, slice_thickness=2.0, roi=128,
                               add_water=False)
    quats = random_uniform_quaternions(3, rng=np.random.default_rng(311))
    mols = [synth.rotated(synth.protein_cloud(n, 310 + i), quats[i])
            for i, n in enumerate((4000, 7000, 2500))]
    for i, m in enumerate(mols):
        m.label = f"mol{i}"
    bad = atm.AtomList(elements=np.array([6, 120]), coordinates=np.zeros((2, 3), np.float32),
                       label="bad")
    got = []
    sim = emsim_amd.simulator.EMSim(pipe, mols[:2] + [bad] + mols[2:],
                                    lambda img, label: got.append((label, img)))
    sim.run()
    assert [label for label, _ in got] == ["mol0", "mol1", "mol2"]
    for (label, img), mol in zip(got, mols):
        assert isinstance(img, np.ndarray) and img.dtype == np.float32
        i32, i64, _, _ = pipe_oracle(mol, tem, 2.0, 2.0, (128, 128))
        e32, e64 = orc.rel_l2(img, i32), orc.rel_l2(img, i64)
        print(f"[EMSim {label}] image vs fftw32 {e32:.2e}, vs fp64 {e64:.2e}", flush=True)
        assert e32 < TOL and e64 < TOL


def _port():
    port = pytest.importorskip("oracle.port")
    try:
        port.load()
    except Exception as exc:
        pytest.fail(f"oracle/fftw_port is not built: {exc}")
    return port


def test_mixed_sizes_full_workload_matches_port():
    """64 molecules of 5k-200k atoms on 128^2 / 256^2 / 512^2 waves in one
    MixedBatch run: one molecule per wave size plus the largest molecule on a
    128^2 wave against oracle/fftw_port.c; six molecules re-simulated alone
    give the same images bit for bit."""
    port = _port()
    tem = _tem()
    items = synth.mixed_workload(count=64, seed=2, lo=5000, hi=200000)
    sizes = np.array([len(m) for m, _ in items])
    print(f"[C4] 64 molecules, {sizes.min()}-{sizes.max()} atoms (mean {sizes.mean():.0f})",
          flush=True)
    mb = MixedBatch(tem, 1.0, 2.0, ELEMS)
    images = mb.run(items)
    torch.cuda.synchronize()
    assert {n for n, _, _ in mb.last_groups} == {128, 256, 512}
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    psi0 = float(np.float32(np.sqrt(tem.electron_dose)))
    checks = []
    for n, ns, idx in mb.last_groups:
        # the group's largest molecule (the 128^2 one is the whole batch's
        # largest on that size: dense slices, atoms outside the window)
        checks.append((n, ns, max(idx, key=lambda i: sizes[i])))
        if n == 512:
            checks.append((n, ns, min(idx, key=lambda i: sizes[i])))
    for n, ns, i in checks:
        mol, _, counts = atm.sort_elements_and_count(atm.centralize(items[i][0]), ELEMS)
        sf = orc.scattering_factors_half(ELEMS, n, n, 1.0, 1.0)
        ref = port.simulate(mol.coordinates, counts, ns, n, 2.0, 1.0, sf, lam, gamma,
                            tem.cs_mm, tem.defocus, tem.aperture, psi0)
        e = orc.rel_l2(images[i].cpu().numpy(), ref)
        print(f"[C4] molecule {i} ({sizes[i]} atoms) on {n}^2 x {ns} slices: image rel L2 vs "
              f"fftw_port {e:.2e}", flush=True)
        assert e < TOL
    # batch invariance
    for n, ns, idx in mb.last_groups:
        pipe = mb.pipe(n, ns)
        for i in idx[:2]:
            alone = pipe.run(pipe.upload([pipe.prepare(atm.centralize(items[i][0]))]))[0]
            assert torch.equal(alone, images[i]), (n, i)


def test_batch_pipe_water_matches_pipe_and_oracle():
    """The reference's default image, Pipe(add_water=True): the batched fused
    path with the same water seed equals Pipe.run, and each image equals the
    oracle chain run on the water-filled histogram the reference-shaped
    builder makes with that seed (binning + add_water on the GPU, whose draw
    statistics are tested separately)."""
    tem = _tem()
    quats = random_uniform_quaternions(2, rng=np.random.default_rng(321))
    mols = [synth.rotated(synth.protein_cloud(n, 320 + i), quats[i])
            for i, n in enumerate((8000, 5000))]
    dz, pixel, n = 2.0, 1.0, 128
    cen = [atm.centralize(m) for m in mols]
    # Pipe sizes the box from molecule 0 (the reference ignores n_slices);
    # molecule 1 is thinner and shares that box
    ns = int(atm.determine_box_size((float(cen[0].space[0]), 1.0, 1.0), (dz, pixel, pixel),
                                    (None, n, n))[0])
    assert float(cen[1].space[0]) <= float(cen[0].space[0])
    np.random.seed(7)
    seeds = [int(np.random.randint(0, 2 ** 63 - 1, dtype=np.int64)) for _ in mols]
    bp = BatchPipe(tem, pixel, dz, n, ns, ELEMS, add_water=True)
    batch = bp.upload([bp.prepare(m) for m in mols])
    images = _host(bp.run(batch, seeds=seeds))
    unfused = _host(bp.run_unfused(batch, seeds=seeds))
    pipe = emsim_amd.pipe.Pipe(tem, resolution=2 * pixel, slice_thickness=dz, roi=n,
                               add_water=True)
    np.random.seed(7)
    via_pipe = _host(pipe.run(mols[0]))
    e_pipe = orc.rel_l2(images[0], via_pipe)
    print(f"[water] fused vs Pipe.run(add_water=True) {e_pipe:.2e}", flush=True)
    assert e_pipe < TOL
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    sf = orc.scattering_factors_half(ELEMS, n, n, pixel, pixel)
    for b in range(2):
        lo, hi = batch.mol_off[b], batch.mol_off[b + 1]
        hist = bp.builder.backend.bin_atoms(batch.coords[lo:hi], batch.counts[b])
        hist = _host(bp.builder.add_water(hist, seed=seeds[b]))
        water_atoms = hist.sum() - (hi - lo)
        v32 = orc.make_multi_slices_fftw32(hist, sf).astype(np.float32)
        w0 = orc.init_wave(n, n, pixel, pixel, tem.electron_dose)
        x32 = orc.multislice_fftw32(w0, v32, dz, lam, gamma, pixel, pixel)
        i32 = orc.intensity(orc.lens_propagate_fftw32(x32, lam, tem.cs_mm, tem.defocus,
                                                      tem.aperture, pixel, pixel))
        e, eu = orc.rel_l2(images[b], i32), orc.rel_l2(unfused[b], i32)
        print(f"[water mol {b}] {water_atoms:.0f} water atoms; fused vs oracle {e:.2e}, "
              f"reference-shaped vs oracle {eu:.2e}", flush=True)
        assert water_atoms > 0.05 * ns * n * n * 0.093
        assert e < TOL and eu < TOL
    # a different seed gives a different image
    other = _host(bp.run(batch, seeds=[seeds[0] + 1, seeds[1]]))
    assert orc.rel_l2(other[0], images[0]) > 1e-3
    np.testing.assert_array_equal(other[1], images[1])


def test_device_preparation_equals_host_preparation():
    """Molecules uploaded as they come (not centred, atoms in random order,
    run_raw) give the images of the host-prepared copies (centralize + sort
    by element, run) bit for bit, alone and through MixedBatch; an atom of an
    element outside the pipe's list is refused on the host."""
    tem = _tem()
    quats = random_uniform_quaternions(4, rng=np.random.default_rng(331))
    rng = np.random.default_rng(332)
    mols = []
    for i, n in enumerate((9000, 3000, 15000, 6000)):
        m = synth.rotated(synth.protein_cloud(n, 330 + i), quats[i])
        perm = rng.permutation(len(m))                       # any atom order
        shift = rng.uniform(-40, 40, 3).astype(np.float32)   # not centred
        mols.append(atm.AtomList(m.elements[perm], m.coordinates[perm] + shift))
    bp = BatchPipe(tem, 1.0, 2.0, 128, 20, ELEMS)
    host = bp.run(bp.upload([bp.prepare(m) for m in mols]))
    dev = bp.run_raw(bp.upload_raw(mols))
    assert torch.equal(host, dev)
    items = [(m, n) for m, n in zip(mols, (128, 256, 128, 256))]
    mb = MixedBatch(tem, 1.0, 2.0, ELEMS)
    a = mb.run(items, device_prep=True)
    b = mb.run(items, device_prep=False)
    assert all(torch.equal(x, y) for x, y in zip(a, b))
    bad = atm.AtomList(np.array([6, 79]), np.zeros((2, 3), np.float32))
    with pytest.raises(ValueError, match="outside"):
        bp.upload_raw([mols[0], bad])


def test_raw_staging_reused_across_calls():
    """Raw uploads go through pinned staging buffers that the next upload
    refills: two batches uploaded back to back before either runs, a larger
    batch after a smaller one (the buffers grow), and MixedBatch run twice in
    a row (its per-size buffers reused) all give the images of separate,
    fully synchronised runs."""
    tem = _tem()
    quats = random_uniform_quaternions(5, rng=np.random.default_rng(341))
    mols = [synth.rotated(synth.protein_cloud(n, 340 + i), quats[i])
            for i, n in enumerate((4000, 7000, 2500, 12000, 5000))]
    bp = BatchPipe(tem, 1.0, 2.0, 128, 16, ELEMS)
    ref = [bp.run(bp.upload([bp.prepare(m)]))[0].clone() for m in mols]
    torch.cuda.synchronize()
    r1 = bp.upload_raw(mols[:2])
    r2 = bp.upload_raw(mols[2:4])          # refills the staging r1 came from
    big = bp.upload_raw(mols)              # grows the buffers
    out1, out2, out3 = bp.run_raw(r1), bp.run_raw(r2), bp.run_raw(big)
    for j in range(2):
        assert torch.equal(out1[j], ref[j]) and torch.equal(out2[j], ref[2 + j])
    for j in range(5):
        assert torch.equal(out3[j], ref[j])
    items = [(m, n) for m, n in zip(mols, (128, 256, 128, 256, 128))]
    mb = MixedBatch(tem, 1.0, 2.0, ELEMS)
    first = [x.clone() for x in mb.run(items)]
    second = mb.run(items)
    assert all(torch.equal(x, y) for x, y in zip(first, second))
