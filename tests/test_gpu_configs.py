"""GPU parity on the large-slab and mixed-size configurations.

Large slab: 2048^2 waves of 0.5 A pixels through water plus a protein-like
cloud, 1 A slices -- here a 6 A thick cut (6 slices, ~0.7M atoms) so the
CPU oracle chain finishes in seconds; the full 400-slice slab uses the same
kernels with more slices.  Mixed sizes: molecules needing 128^2, 256^2 and
512^2 waves, batched per size; each image must equal the same molecule
simulated alone (the fused path is batch-invariant) and match the oracle.
"""
import numpy as np
import pytest
import torch

import emsim_amd
from emsim_amd import atoms as atm
from emsim_amd import synth
from emsim_amd.batch import BatchPipe, MixedBatch
from oracle import emsim_oracle as orc
from tests._chain import slices_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5
ELEMS = [1, 6, 7, 8, 16]


def _tem():
    return emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                             aperture=10e-3)


def _oracle(mol, tem, n, ns, dz, pixel):
    mol = atm.centralize(mol)
    v32, v64, _ = slices_oracle(mol, dz, pixel, (n, n), ns)
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    w0 = orc.init_wave(n, n, pixel, pixel, tem.electron_dose)
    x32 = orc.multislice_fftw32(w0, v32.astype(np.float32), dz, lam, gamma, pixel, pixel)
    del v32
    x64 = orc.multislice_f64(w0, v64, dz, lam, gamma, pixel, pixel)
    del v64
    i32 = orc.intensity(orc.lens_propagate_fftw32(x32, lam, tem.cs_mm, tem.defocus,
                                                  tem.aperture, pixel, pixel))
    return x32, x64, i32


@pytest.fixture(autouse=True)
def _gpu(hip):
    yield
    torch.cuda.synchronize()


def test_large_slab_cut_2048_matches_oracle():
    tem = _tem()
    n, ns, dz, pixel = 2048, 6, 1.0, 0.5
    water = synth.water_slab(6.0, 1024.0, 1024.0, seed=1)
    cloud = synth.protein_cloud(50000, 2001)
    keep = np.abs(cloud.coordinates[:, 0]) < 2.9
    cut = atm.AtomList(cloud.elements[keep], cloud.coordinates[keep])
    mol = atm.concatenate([water, cut])
    print(f"[large slab cut] {len(mol)} atoms, {n}^2, {ns} slices", flush=True)
    pipe = BatchPipe(tem, pixel, dz, n, ns, ELEMS)
    batch = pipe.upload([pipe.prepare(mol)])
    wave = pipe.exit_waves(batch)[0].cpu().numpy()
    image = pipe.run(batch)[0].cpu().numpy()
    x32, x64, i32 = _oracle(mol, tem, n, ns, dz, pixel)
    e32, e64, ei = orc.rel_l2(wave, x32), orc.rel_l2(wave, x64), orc.rel_l2(image, i32)
    print(f"[large slab cut] exit wave {e32:.2e} (fp64 {e64:.2e}), image {ei:.2e}", flush=True)
    assert e32 < TOL and e64 < TOL and ei < TOL


def test_mixed_sizes_batch_invariant_and_match_oracle():
    tem = _tem()
    items = synth.mixed_workload(count=9, seed=2, lo=1500, hi=6000)
    mb = MixedBatch(tem, 1.0, 2.0, ELEMS)
    images = mb.run(items)
    assert len(mb.last_groups) >= 2
    for n, ns, idx in mb.last_groups:
        print(f"[mixed] size {n}: {len(idx)} molecules, {ns} slices", flush=True)
        pipe = mb.pipe(n, ns)
        for i in idx:
            alone = pipe.run(pipe.upload([pipe.prepare(atm.centralize(items[i][0]))]))[0]
            assert torch.equal(alone, images[i])
    # two molecules of different sizes against the oracle
    checked = 0
    for n, ns, idx in mb.last_groups[:2]:
        i = idx[0]
        _, _, i32 = _oracle(items[i][0], tem, n, ns, 2.0, 1.0)
        err = orc.rel_l2(images[i].cpu().numpy(), i32)
        print(f"[mixed] size {n} molecule {i}: image {err:.2e}", flush=True)
        assert err < TOL
        checked += 1
    assert checked == 2


def test_large_slab_full_400_slices_matches_oracle_fixture():
    """The large slab at its full depth: 2048^2, 0.5 A pixels, all 400 slices of
    1 A through 42.3M atoms (water plus five 50k-atom clouds), against the CPU
    oracle's result at the same configuration (tests/golden/make_c3_fixture.py
    ran oracle/fftw_port.c on the same synthetic atoms): image and exit-wave
    crops around the centre, 32 x 32 block sums of the whole image, and the
    image's total and sum of squares.  This is where float32 rounding
    accumulates over the most slices."""
    from pathlib import Path
    from tests.golden import make_c3_fixture as fx_mod
    fx = np.load(Path(fx_mod.__file__).with_name("c3_slab_2048x400.npz"))
    tem = _tem()
    mol = synth.large_slab(seed=1)
    pipe = BatchPipe(tem, fx_mod.PIXEL, fx_mod.DZ, fx_mod.N, fx_mod.NS, ELEMS)
    prep = pipe.prepare(mol)
    assert len(mol) == int(fx["n_atoms"][0])
    assert np.array_equal(prep.counts, fx["counts"])
    print(f"[large slab 400] {len(mol)} atoms, {fx_mod.N}^2 x {fx_mod.NS} slices", flush=True)
    batch = pipe.upload([prep])
    del mol, prep
    wave = pipe.exit_waves(batch)[0].cpu().numpy()
    image = pipe.run(batch)[0].cpu().numpy()
    got = fx_mod.summarize(image, wave)
    errs = {k: orc.rel_l2(got[k], fx[k]) for k in ("image_crop", "wave_crop", "block_sums")}
    for k in ("image_sum", "image_sumsq"):
        errs[k] = abs(float(got[k][0]) - float(fx[k][0])) / abs(float(fx[k][0]))
    print("[large slab 400] rel. differences vs oracle: " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < TOL for v in errs.values()), errs
