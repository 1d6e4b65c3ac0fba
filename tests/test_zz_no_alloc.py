"""emsim_hip_orient_batch_noalloc (no workspace argument, nothing allocated
by the library) against emsim_hip_orient_batch with a caller workspace: the
same kernels, so the coordinates must agree bit for bit.  Two calls in a
row, the second with a smaller molecule and batch, so the library's
scratch holds the first call's bounding boxes when the second starts."""
import numpy as np
import pytest
import torch

from emsim_amd import synth
from emsim_amd.backend._device import ptr, stream_handle
from emsim_amd.batch import BatchPipe
from emsim_amd.tem import TEM
from emsim_amd.utils.rot import get_rotation_mattrices, random_uniform_quaternions

pytestmark = pytest.mark.gpu

ELEMS = [1, 6, 7, 8, 16]


def _inputs(n_atoms, batch, seed):
    tem = TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700, aperture=10e-3)
    pipe = BatchPipe(tem, 1.0, 2.0, 64, 4, ELEMS)
    base = pipe.prepare(synth.protein_cloud(n_atoms, seed))
    c = base.coords
    centre = c.min(axis=0) + (c.max(axis=0) - c.min(axis=0)) / 2.0
    quats = random_uniform_quaternions(batch, rng=np.random.default_rng(seed))
    rot = get_rotation_mattrices(quats).astype(np.float32).reshape(batch, 9)
    dev = torch.device("cuda", torch.cuda.current_device())
    return (torch.as_tensor(c).to(dev), torch.as_tensor(rot).to(dev), centre.astype(np.float64),
            c.shape[0])


def _with_workspace(hip, base, rot, centre, n, batch):
    out = torch.empty((batch * n, 3), dtype=torch.float32, device=base.device)
    ws = torch.empty(int(hip.raw("emsim_hip_orient_workspace_size")(batch)), dtype=torch.uint8,
                     device=base.device)
    hip.call("emsim_hip_orient_batch", ptr(base), n, ptr(rot), float(centre[0]),
             float(centre[1]), float(centre[2]), batch, ptr(out), ptr(ws), ws.numel(),
             stream_handle())
    return out


def _noalloc(hip, base, rot, centre, n, batch):
    out = torch.empty((batch * n, 3), dtype=torch.float32, device=base.device)
    hip.call("emsim_hip_orient_batch_noalloc", ptr(base), n, ptr(rot), float(centre[0]),
             float(centre[1]), float(centre[2]), batch, ptr(out), stream_handle())
    return out


def test_noalloc_orientations_match_workspace_entry_point(hip):
    for n_atoms, batch, seed in ((20000, 64, 11), (3000, 7, 12)):
        print(f"orient_batch_noalloc: {n_atoms} atoms x {batch} orientations, seed {seed}",
              flush=True)
        base, rot, centre, n = _inputs(n_atoms, batch, seed)
        got = _noalloc(hip, base, rot, centre, n, batch)
        want = _with_workspace(hip, base, rot, centre, n, batch)
        torch.cuda.synchronize()
        assert torch.isfinite(got).all().item()
        assert torch.equal(got, want)
        # every instance is centred: its bounding box is symmetric about 0
        g = got.reshape(batch, n, 3)
        mid = (g.amax(dim=1) + g.amin(dim=1)) / 2
        assert mid.abs().max().item() < 1e-3


def test_noalloc_rejects_batches_beyond_its_scratch(hip):
    limit = 4096
    base, rot, centre, n = _inputs(100, 1, 13)
    out = torch.empty((n, 3), dtype=torch.float32, device=base.device)
    rc = hip.raw("emsim_hip_orient_batch_noalloc")(ptr(base), n, ptr(rot), 0.0, 0.0, 0.0,
                                                   limit + 1, ptr(out), stream_handle())
    assert rc == hip.EUNSUPPORTED
    assert "scratch" in hip.last_error()
