"""Train the learned coarse-slice correction on pairs the plain GPU path makes.

    python -m emsim_amd.train_fno --steps 2500 --batch 8 --coarse-dz 25 --out fno_512.npz

Every step draws a batch of synthetic protein-like clouds (atom counts
log-uniform in 10k-100k) from a pool of base clouds (seeds 101 and up) in
fresh random orientations, simulates them twice with the plain fused
multislice kernels -- at the coarse slice thickness and at 1 A -- and takes
one Adam step of the network on (coarse exit wave, coarse potentials) ->
fine exit wave.  At the end the network is evaluated on held-out clouds
(seeds 10001 and up, disjoint from the pool and from the benchmark's
20001-20064) against plain 1 A multislice, with and without the correction.
"""
import argparse
import json
import time

import numpy as np
import torch

from . import synth
from .learned import FNO, CoarseFinePairs, train, wave_loss
from .tem import TEM
from .utils.rot import random_uniform_quaternions


def default_tem() -> TEM:
    return TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700, aperture=10e-3)


def cloud_pool(count: int, seed: int, lo: int = 10000, hi: int = 100000):
    """count base clouds, atom counts log-uniform in [lo, hi], cloud seeds
    seed + 1 ... seed + count (built on a thread pool)."""
    from .batch import pool_map
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(int)
    return pool_map(lambda i: synth.protein_cloud(int(sizes[i]), seed + 1 + i), range(count))


def batch_source(pipe, pool, batch: int, seed: int):
    """batches(step): batch random (cloud, orientation) draws, prepared."""
    def batches(step: int):
        rng = np.random.default_rng([seed, step])
        idx = rng.integers(0, len(pool), batch)
        quats = random_uniform_quaternions(batch, rng=rng)
        from .batch import pool_map
        return pool_map(lambda k: pipe.prepare(synth.rotated(pool[idx[k]], quats[k])),
                        range(batch))
    return batches


def evaluate(fno: FNO, pairs: CoarseFinePairs, mols) -> dict:
    """Mean / max exit-wave relative L2 against plain fine-slice multislice."""
    pb = pairs.fine.upload([pairs.fine.prepare(m) for m in mols])
    cw, cs, fw = pairs.make(pb)
    x = fno.pack(cw, cs, pairs.psi0, pairs.sigma)
    pred = fno.apply(cw, fno.forward(x), pairs.psi0)
    learned, _ = wave_loss(pred, fw)
    coarse, _ = wave_loss(cw, fw)
    l = learned[:, 0].sqrt().cpu().numpy()
    c = coarse[:, 0].sqrt().cpu().numpy()
    return {"images": len(mols), "learned_mean": float(l.mean()), "learned_max": float(l.max()),
            "coarse_mean": float(c.mean()), "coarse_max": float(c.max())}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--steps", type=int, default=2500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--slices", type=int, default=100)
    ap.add_argument("--coarse-dz", type=float, default=25.0)
    ap.add_argument("--channels", type=int, default=32)
    ap.add_argument("--modes", type=int, default=16)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pool", type=int, default=256, help="base training clouds (seeds 101+)")
    ap.add_argument("--eval", type=int, default=32, help="held-out evaluation images")
    ap.add_argument("--init", default=None, help="checkpoint to continue from")
    ap.add_argument("--out", required=True)
    ap.add_argument("--log", default=None, help="JSON history file")
    a = ap.parse_args(argv)

    tem = default_tem()
    coarse_slices = int(np.ceil(a.slices * 1.0 / a.coarse_dz))
    pairs = CoarseFinePairs(tem, n=a.n, fine_slices=a.slices, coarse_dz=a.coarse_dz,
                            coarse_slices=coarse_slices)
    if a.pool > 9000:
        ap.error("--pool must stay below 9000 (seeds 10001+ are the held-out clouds)")
    if a.init:
        fno = FNO.load(a.init)
    else:
        fno = FNO(a.n, 2 + coarse_slices, a.channels, a.modes, a.layers)
        fno.set_params(fno.init_params(seed=100))
    t0 = time.perf_counter()
    pool = cloud_pool(a.pool, seed=100)
    print(f"training pool: {len(pool)} clouds in {time.perf_counter() - t0:.0f} s", flush=True)
    hist = train(fno, pairs, batch_source(pairs.fine, pool, a.batch, seed=100), a.steps,
                 lr=a.lr, log_every=max(a.steps // 25, 1),
                 log=lambda s: print(s, flush=True))
    fno.save(a.out, coarse_dz=a.coarse_dz)
    eval_pool = cloud_pool(a.eval, seed=10000)     # one distinct cloud per image
    quats = random_uniform_quaternions(a.eval, rng=np.random.default_rng(2))
    mols = [synth.rotated(eval_pool[i], quats[i]) for i in range(a.eval)]
    ev = evaluate(fno, pairs, mols)
    print("held-out:", json.dumps(ev), flush=True)
    if a.log:
        with open(a.log, "w") as fh:
            json.dump({"args": vars(a), "history": hist, "eval": ev}, fh, indent=1)


if __name__ == "__main__":
    main()
