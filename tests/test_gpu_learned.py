"""GPU tests of the learned coarse-slice pipeline: what the trained network
gains over coarse slicing alone (held-out molecules, judged by plain 1 A
multislice), and the gate that falls back to plain fine-slice multislice."""
from pathlib import Path

import numpy as np
import pytest
import torch

import emsim_amd
from emsim_amd import synth
from emsim_amd.utils.rot import random_uniform_quaternions

pytestmark = pytest.mark.gpu

CKPT = Path(__file__).resolve().parent.parent / "emsim_amd" / "data" / "fno_512.npz"


def _tem():
    return emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                             aperture=10e-3)


@pytest.fixture(scope="module")
def setup(hip):
    if not CKPT.exists():
        pytest.skip("no trained network (python -m emsim_amd.train_fno --out "
                    "emsim_amd/data/fno_512.npz)")
    from emsim_amd.learned import FNO, CoarseFinePairs
    fno = FNO.load(str(CKPT))
    coarse_dz = fno.coarse_dz if fno.coarse_dz is not None else 8.0
    pairs = CoarseFinePairs(_tem(), n=fno.n, fine_slices=100, coarse_dz=coarse_dz,
                            coarse_slices=fno.cin - 2)
    # held out: seeds far from the training pool's (101 ...)
    sizes = [12000, 25000, 40000, 70000]
    quats = random_uniform_quaternions(len(sizes), rng=np.random.default_rng(20001))
    mols = [synth.rotated(synth.protein_cloud(s, 20000 + i), quats[i]) for i, s in enumerate(sizes)]
    batch = pairs.fine.upload([pairs.fine.prepare(m) for m in mols])
    return fno, pairs, batch


def test_network_gain_over_coarse_slicing(setup):
    from emsim_amd.learned import wave_loss
    fno, pairs, batch = setup
    cw, cs, fine = pairs.make(batch)
    x = fno.pack(cw, cs, pairs.psi0, pairs.sigma)
    pred = fno.apply(cw, fno.forward(x), pairs.psi0)
    learned = wave_loss(pred, fine)[0][:, 0].sqrt().cpu().numpy()
    coarse = wave_loss(cw, fine)[0][:, 0].sqrt().cpu().numpy()
    print(f"exit-wave rel L2 vs plain 1 A multislice: coarse slices alone {coarse.round(5)}, "
          f"with the network {learned.round(5)}")
    assert learned.max() < 0.01
    assert learned.mean() < 0.5 * coarse.mean()
    assert (learned < coarse).all()


def test_gate_passes_and_matches_components(setup):
    from emsim_amd.learned import LearnedPipe
    fno, pairs, batch = setup
    lp = LearnedPipe(fno, pairs)
    wave, rep = lp.exit_waves(batch)
    torch.cuda.synchronize()
    assert not rep.fallback, rep.reason
    assert rep.sampled == 0 and rep.sampled_rel_l2 < 0.01
    assert max(rep.norm_dev) < 0.01
    ref, _ = lp.exit_waves(batch, check=False)
    assert torch.equal(wave, ref)


@pytest.mark.parametrize("which", ["norm", "sampled"])
def test_gate_failure_falls_back_to_plain_multislice(setup, which):
    from emsim_amd.learned import LearnedPipe
    fno, pairs, batch = setup
    kw = {"norm_tol": 0.0} if which == "norm" else {"check_tol": 0.0}
    lp = LearnedPipe(fno, pairs, **kw)
    wave, rep = lp.exit_waves(batch)
    assert rep.fallback and rep.reason
    assert torch.equal(wave, pairs.fine.exit_waves(batch))
    img, rep2 = lp.run(batch)
    assert rep2.fallback and img.shape == (len(batch), fno.n, fno.n)
