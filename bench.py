"""Throughput benchmark of the hip multislice path on one MI355X.

Workload (the headline metric's configuration): 512 x 512 waves, 1 A pixel,
100 slices of 1 A, 200 keV, Cs 1.3 mm, defocus 700 A, aperture 10 mrad,
64 synthetic protein-like clouds (H/C/N/O/S, 1 A minimum spacing), each
drawn independently with a log-uniform atom count in [10k, 100k] and a
uniformly random orientation, no water.  Their seeds (20001-20064) are
disjoint from the correction network's training clouds (101-132) and its
held-out evaluation clouds (10001-10008).  One step is one pass of the full
path over the batch, atoms already in GPU memory: binning -> potential
slices -> multislice -> lens -> |psi|^2 (emsim_amd.batch.BatchPipe).

Prints one JSON line.  ``value`` comes from the timed steps run without
instrumentation.  The same number of steps is then run again with every
launch bracketed by HIP events on its stream (the per-kernel table).
``roofline`` describes the kernel with the largest share of GPU time there:
``achieved`` is the bytes that kernel must move (compulsory_bytes) divided by
its mean launch duration, against the 8 TB/s HBM peak; ``traffic`` and
``measured_frac`` come from the committed rocprofv3 FETCH/WRITE summary and
``valu_issue_frac`` from the committed SQ instruction-counter summary, each
used only when it was measured on this exact library.  ``cpu_baseline``
times the CPU restatement on a bounded sample of the same workload;
``cpu_rows`` times the numpy and C/OpenMP restatements on the other
configurations; ``end_to_end`` runs the batched EMSim loop from host atoms
to host images.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(HERE, "profiles")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--slices", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-rows", action="store_true",
                    help="skip the CPU restatements of the other configurations")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the host-atoms -> host-images (BatchSim) line")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (C2 orientations, C3 slab, C4 mixed)")
    ap.add_argument("--unfused", action="store_true",
                    help="time the reference-shaped kernels (histogram, make_slices, multislice)")
    ap.add_argument("--no-timing", action="store_true",
                    help="do not bracket launches with HIP events")
    ap.add_argument("--traffic", default=os.path.join(PROFILES, "traffic_512x100_b64.json"),
                    help="rocprofv3 FETCH/WRITE summary (emsim_amd.profiling)")
    ap.add_argument("--valu", default=os.path.join(PROFILES, "valu_512x100_b64.json"),
                    help="rocprofv3 SQ instruction-counter summary (emsim_amd.profiling --sq)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 4 device-to-device copies of MIB MiB (a known-size "
                         "reference for HBM byte counters) and exit")
    ap.add_argument("--learned", default=os.path.join(HERE, "emsim_amd", "data", "fno_512.npz"),
                    help="trained correction network for the learned coarse-slice line")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-only", action="store_true",
                    help="run only the learned coarse-slice line (prints its JSON object)")
    ap.add_argument("--config-only", choices=["C2", "C3", "C4"], default=None,
                    help="run only one configuration line (prints its JSON object)")
    return ap.parse_args()


# ---------------------------------------------------------------------------
# workload
# ---------------------------------------------------------------------------
def workload(batch: int, n_atoms_range=(10000, 100000), seed: int = 20000):
    """batch independent clouds: atom counts log-uniform in n_atoms_range,
    cloud seeds seed + 1 ... seed + batch, uniform random orientations."""
    from emsim_amd import synth
    from emsim_amd.batch import pool_map
    from emsim_amd.utils.rot import random_uniform_quaternions
    rng = np.random.default_rng(seed)
    lo, hi = np.log(n_atoms_range[0]), np.log(n_atoms_range[1])
    sizes = np.exp(rng.uniform(lo, hi, batch)).astype(int)
    quats = random_uniform_quaternions(batch, rng=rng)
    mols = pool_map(lambda i: synth.rotated(synth.protein_cloud(int(sizes[i]), seed + 1 + i),
                                            quats[i]), range(batch))
    return mols, sizes


def live_columns(n: int, d: float) -> int:
    """Wave columns the 2/3 propagator mask keeps (fy^2 <= filter^2 in
    float32, the test the library makes on the host); the others stay zero
    between slices and are never read or written by the passes."""
    f32 = np.float32
    dfy = f32(f32(1.0) / f32(d)) / f32(n)
    filt = f32(f32(0.6667) * f32(f32(0.5) / f32(d)))
    k = np.arange(n)
    k = np.where(k < n // 2, k, k - n).astype(f32)
    fy = k * dfy
    return int(np.count_nonzero(fy * fy <= filt * filt))


def compulsory_bytes(name: str, ctx) -> float:
    """HBM bytes one launch of `name` must move in this workload (mean over
    its launches), at the data types the kernels use.

    With B molecules, n x n waves, H = n/2 + 1 half-spectrum columns, nE
    elements, nS slices and L live wave columns:
      k_spec_sparse          Y write B n H 8 + form factors nE n H 4
                             + row table B nE n 4 + atom y bins 2 per binned atom
      k_slice_pass_a         Y read B n H 8 + wave write B n L 8
                             + wave read B n L 8 (not on the first slice)
      k_col_fft_mul_ifft     wave read + write B n L 16 (last slice: all n columns)
      k_row_ifft_intensity   wave read B n n 8 + image write B n n 4
      k_csr_classify         coordinates 12 + row 4 + y bin 2 per atom
      k_csr_scatter          row 4 + y bin 2 read + bucket 2 write per atom
      k_fill_*               the filled buffer
    The reference-shaped kernels (--unfused) are charged their dense passes:
      k_row_r2c_pairs        histogram 4 nE + half spectra 8h nE per pixel and slice
      k_col_spectrum         half spectra 8h nE in + 8h out (+ form factors)
      k_row_c2r_pairs        half spectrum 8h + potential 4
      k_bin_tiles            histogram 4 nE
    """
    B, n, nS, nE, L = ctx["batch"], ctx["n"], ctx["slices"], ctx["n_elems"], ctx["live"]
    H = n // 2 + 1
    h = H / n
    npix = n * n
    atoms = ctx["n_atoms"]
    per = {
        "k_spec_sparse": B * n * H * 8.0 + nE * n * H * 4.0 + B * nE * n * 4.0
        + 2.0 * ctx["binned_atoms"] / nS,
        "k_slice_pass_a": B * n * H * 8.0 + B * n * L * 8.0 + B * n * L * 8.0 * (nS - 1) / nS,
        "k_col_fft_mul_ifft": B * n * 16.0 * (L * (nS - 1) + n) / nS,
        "k_row_ifft_intensity": 12.0 * npix * B,
        "k_csr_classify": 18.0 * atoms,
        "k_csr_scatter": 8.0 * atoms,
        "k_fill_complex": 8.0 * npix * B,
        "k_row_transmit_fft": 20.0 * npix * B,
        "k_row_fft": 16.0 * npix * B,
        "k_row_r2c_pairs": (4.0 * nE + 8 * h * nE) * npix * nS,
        "k_col_spectrum": (8 * h * nE + 8 * h) * npix * nS + 4 * h * nE * npix,
        "k_row_c2r_pairs": (8 * h + 4.0) * npix * nS,
        "k_bin_tiles": 4.0 * nE * npix * nS,
    }
    for k, v in per.items():
        if name.startswith(k):
            return v
    return float("nan")


def dense_reference_bytes_per_image(n: int, nS: int, nE: int) -> float:
    """The reference's dense passes per image at 4 B histograms and 8h B half
    spectra: nS (nE (8 + 16h) + 16h + 4 + 36) + 44 bytes per pixel."""
    h = (n // 2 + 1) / n
    return (nS * (nE * (8 + 16 * h) + 16 * h + 4 + 36) + 44) * n * n


def _profile(path: str):
    """A committed profile summary, if it describes the library in use."""
    try:
        from emsim_amd.profiling import library_sha256
        data = json.loads(open(path).read())
        if data["library_sha256"] != library_sha256():
            return None, "profile measured on another build of the library"
        return data, None
    except FileNotFoundError:
        return None, "no profile committed"
    except Exception as e:   # pragma: no cover - malformed file
        return None, f"profile unreadable: {e}"


def _lookup(data, kernel: str):
    base = kernel.split("<")[0]
    for name, rec in data["kernels"].items():
        if name.split("<")[0].split("(")[0].strip() == base:
            return rec
    return None


# ---------------------------------------------------------------------------
# learned line
# ---------------------------------------------------------------------------
def fno_flops_per_image(n: int, cin_pad: int, C: int, modes: int, layers: int) -> float:
    """2 M N K over the network's GEMMs for one n x n image (truncated DFTs
    counted as GEMMs against their bases; the per-mode channel mix adds
    8 (2 modes) modes C^2)."""
    M1, M2 = 2 * modes, modes
    layer = (2.0 * (2 * M2) * C * n * n            # y-transform, per row
             + 2.0 * (2 * M1) * (2 * M2 * C) * n   # x-transform
             + 8.0 * M1 * M2 * C * C               # channel mix
             + 2.0 * (2 * n) * (M2 * C) * (2 * M1) # inverse x-transform
             + 2.0 * n * n * C * (C + 2 * M2))     # W v + inverse y-transform
    return 2.0 * n * n * cin_pad * C + layers * layer + 2.0 * n * n * C * 2


def learned_bench(args, cfg, tem, mols, plain_ips):
    """Coarse fused multislice + FNO correction + lens, gated, on the same
    molecules; error against plain fine-slice multislice."""
    import torch
    from emsim_amd.backend import hip_ext
    from emsim_amd.learned import FNO, CoarseFinePairs, LearnedPipe, wave_loss

    fno = FNO.load(args.learned)
    coarse_slices = fno.cin - 2
    coarse_dz = fno.coarse_dz if fno.coarse_dz is not None else 8.0
    pairs = CoarseFinePairs(tem, n=cfg["n"], pixel=cfg["pixel"], fine_dz=cfg["dz"],
                            fine_slices=cfg["slices"], coarse_dz=coarse_dz,
                            coarse_slices=coarse_slices, elements=cfg["elements"])
    lp = LearnedPipe(fno, pairs)
    batch = pairs.fine.upload(pairs.fine.prepare_many(mols))
    for _ in range(args.warmup):
        lp.run(batch)
    torch.cuda.synchronize()
    reports = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        img, rep = lp.run(batch)
        reports.append(rep)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # kernel table from a second, event-instrumented pass of the same steps
    kernels, elapsed_instr = [], None
    if not args.no_timing:
        hip_ext.timing_collect()
        hip_ext.timing_enable(True)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            lp.run(batch)
        torch.cuda.synchronize()
        elapsed_instr = time.perf_counter() - t1
        hip_ext.timing_enable(False)
        kernels = hip_ext.timing_collect()
    assert torch.isfinite(img).all().item()
    value = args.batch * args.steps / elapsed
    # the same with the sampled fine-slice check on every 4th batch only
    # (norm check still on every image)
    lp4 = LearnedPipe(fno, pairs, check_every=4)
    lp4.run(batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(args.steps, 4)):
        img4, rep4 = lp4.run(batch)
    torch.cuda.synchronize()
    value4 = args.batch * max(args.steps, 4) / (time.perf_counter() - t0)
    # the coarse simulation alone (no network, no check): what the network costs
    cs = pairs.coarse
    for _ in range(2):
        cs.run(batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cs.run(batch)
    torch.cuda.synchronize()
    coarse_ips = args.batch * args.steps / (time.perf_counter() - t0)
    # accuracy on the benchmark batch (outside the timed region)
    wave, _ = lp.exit_waves(batch, check=False)
    cw, _, fine = pairs.make(batch)
    err = wave_loss(wave, fine)[0][:, 0].sqrt().cpu().numpy()
    base = wave_loss(cw, fine)[0][:, 0].sqrt().cpu().numpy()
    # MFMA kernels: the generic GEMM and the fused row kernels
    gemm_ms = sum(k["total_ms"] for k in kernels
                  if k["name"].startswith(("k_gemm", "k_fno_rows"))) / args.steps
    mix_ms = sum(k["total_ms"] for k in kernels
                 if k["name"].startswith("k_fno") and not k["name"].startswith("k_fno_rows")) \
        / args.steps
    flops = fno_flops_per_image(cfg["n"], fno.cin_pad, fno.C, fno.modes, fno.L) * args.batch
    tflops = flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    kernel_total = sum(k["total_ms"] for k in kernels)
    top = sorted(kernels, key=lambda r: -r["total_ms"])[:10]
    return {
        "images_per_s": round(value, 2),
        "speedup_vs_plain": round(value / plain_ips, 3) if plain_ips else None,
        "images_per_s_fine_check_every_4th_batch": round(value4, 2),
        "speedup_vs_plain_fine_check_every_4th_batch":
            round(value4 / plain_ips, 3) if plain_ips else None,
        "coarse_only_images_per_s": round(coarse_ips, 2),
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "slices_per_s_fine_equivalent": round(value * cfg["slices"], 1),
        "coarse_dz": coarse_dz, "coarse_slices": coarse_slices,
        "network": {"layers": fno.L, "channels": fno.C, "modes": fno.modes,
                    "dtype": "bf16 MFMA, f32 accumulate", "train_steps": fno.step},
        "exit_wave_rel_l2_vs_plain": {"mean": float(err.mean()), "max": float(err.max()),
                                      "coarse_alone_mean": float(base.mean()),
                                      "coarse_alone_max": float(base.max()),
                                      "molecules": "the benchmark batch (held out from "
                                                   "training and from the network's "
                                                   "evaluation set)"},
        "gate": {"checks": "norm conservation per image (1%) + one sampled image per batch "
                           "re-simulated with plain 1 A multislice (1% rel L2)",
                 "fallbacks": sum(1 for r in reports if r.fallback),
                 "sampled_rel_l2": [round(r.sampled_rel_l2, 6) for r in reports]},
        "fno": {"gflop_per_image": round(flops / args.batch / 1e9, 3),
                "mfma_kernels_ms_per_step": round(gemm_ms, 3),
                "other_fno_ms_per_step": round(mix_ms, 3),
                "achieved_tflops": None if tflops is None else round(tflops, 1),
                "mfma_peak_tflops": 2500.0,
                "mfma_frac": None if tflops is None else round(tflops / 2500.0, 4)},
        # the sampled fine check runs on a second stream beside the coarse run
        # and the network, so summed kernel time can exceed the wall time
        "kernel_time_over_wall_two_streams":
            round(kernel_total * 1e-3 / elapsed_instr, 4) if kernels else None,
        "kernels": [{"kernel": k["name"], "launches": k["count"],
                     "mean_us": round(k["mean_ms"] * 1e3, 2),
                     "ms_per_step": round(k["total_ms"] / args.steps, 3)} for k in top],
    }


# ---------------------------------------------------------------------------
# host-to-host line
# ---------------------------------------------------------------------------
def end_to_end(pipe, mols, batch: int, repeats: int = 6):
    """BatchSim (the batched EMSim loop) from host AtomLists to host images:
    centring and sort by element on a host thread pool, one upload per batch,
    the fused GPU path, and a D2H copy into pinned buffers overlapped with the
    next batch, delivered to a callback."""
    from emsim_amd.batch import BatchSim
    got = []
    sim = BatchSim(pipe, list(mols), lambda img, label: got.append(img.shape), batch=batch)
    sim.run()   # warm-up (pinned ring, first-call setup)
    got.clear()
    stream = [m for _ in range(repeats) for m in mols]
    sim = BatchSim(pipe, stream, lambda img, label: got.append(img.shape), batch=batch)
    t0 = time.perf_counter()
    delivered = sim.run()
    dt = time.perf_counter() - t0
    assert delivered == len(stream) == len(got)
    return {"images_per_s": round(delivered / dt, 2), "images": delivered,
            "seconds": round(dt, 3),
            "region": "host AtomLists (unsorted, uncentred) -> host float32 images: "
                      "host preparation on a thread pool, upload, fused GPU path, "
                      "pinned D2H ring overlapped with the next batch"}


# ---------------------------------------------------------------------------
# other configurations on the GPU
# ---------------------------------------------------------------------------
def _busy_frac(fn):
    """Fraction of one call of fn's wall time that its kernels run: summed
    per-launch event intervals over the wall time of that same call (the
    events add a little time to every launch, so the ratio is taken within
    the instrumented call, not against an uninstrumented one)."""
    import torch
    from emsim_amd.backend import hip_ext
    hip_ext.timing_collect()
    torch.cuda.synchronize()
    hip_ext.timing_enable(True)
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    hip_ext.timing_enable(False)
    kernel_ms = sum(k["total_ms"] for k in hip_ext.timing_collect())
    return round(kernel_ms * 1e-3 / wall, 3)


def _wall(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def config_lines(tem, which=("C2", "C4", "C3")):
    """C2 (one 50k-atom cloud in 64 orientations rotated on the device), C3
    (the 2048^2 x 400-slice slab, one image), C4 (256 molecules of 5k-200k
    atoms on 128^2/256^2/512^2 waves, from host AtomLists as they come):
    images/s and the fraction of the wall time the GPU spends in kernels."""
    out = {}
    if "C2" in which:
        out.update(_config_c2(tem))
    if "C4" in which:
        out.update(_config_c4(tem))
    if "C3" in which:
        out.update(_config_c3(tem))
    return out


def _config_c2(tem):
    from emsim_amd import atoms as atm
    from emsim_amd import synth
    from emsim_amd.batch import BatchPipe
    from emsim_amd.utils.rot import random_uniform_quaternions
    out = {}
    mol = synth.protein_cloud(50000, 0)
    ns2 = int(np.ceil(float(atm.centralize(mol).space[0]) / 2.0))
    p2 = BatchPipe(tem, 1.0, 2.0, 256, ns2, [1, 6, 7, 8, 16])
    b2 = p2.upload_orientations(mol, random_uniform_quaternions(64, rng=np.random.default_rng(2)))
    t = _wall(lambda: p2.run(b2), 5)
    busy = _busy_frac(lambda: p2.run(b2))
    one = p2.upload([p2.prepare(mol)])
    t1 = _wall(lambda: p2.run(one), 5)
    out["C2"] = {"workload": f"50k-atom cloud (seed 0) in 64 orientations, 256^2, {ns2} slices "
                             f"of 2 A, 1 A pixels",
                 "images_per_s": round(64 / t, 1), "ms_per_image_batch1": round(t1 * 1e3, 3),
                 "gpu_busy_frac": busy}
    progress(f"C2: {out['C2']['images_per_s']} images/s")
    return out


def _config_c4(tem):
    from emsim_amd import synth
    from emsim_amd.batch import MixedBatch
    out = {}
    t0 = time.perf_counter()
    items = synth.mixed_workload(count=256, seed=2, lo=5000, hi=200000)
    gen = time.perf_counter() - t0
    mb = MixedBatch(tem, 1.0, 2.0, [1, 6, 7, 8, 16])
    t = _wall(lambda: mb.run(items), 2)
    busy = _busy_frac(lambda: mb.run(items))
    th = _wall(lambda: mb.run(items, device_prep=False), 1)
    sizes = np.array([len(m) for m, _ in items])
    out["C4"] = {"workload": f"256 molecules of {sizes.min()}-{sizes.max()} atoms (mean "
                             f"{sizes.mean():.0f}), waves 128^2/256^2/512^2, 1 A pixels, 2 A "
                             f"slices; host AtomLists -> GPU images, preparation on the device",
                 "images_per_s": round(256 / t, 1), "ms_per_256": round(t * 1e3, 1),
                 "gpu_busy_frac": busy,
                 "groups": [(n, ns, len(idx)) for n, ns, idx in mb.last_groups],
                 "host_prepared_images_per_s": round(256 / th, 1),
                 "generation_s": round(gen, 1)}
    progress(f"C4: {out['C4']['images_per_s']} images/s")
    return out


def _config_c3(tem):
    from emsim_amd import synth
    from emsim_amd.batch import BatchPipe
    out = {}
    slab = synth.large_slab(seed=1)
    p3 = BatchPipe(tem, 0.5, 1.0, 2048, 400, [1, 6, 7, 8, 16])
    b3 = p3.upload([p3.prepare(slab)])
    n_atoms = len(slab)
    del slab
    t = _wall(lambda: p3.run(b3), 2)
    busy = _busy_frac(lambda: p3.run(b3))
    out["C3"] = {"workload": f"2048^2 x 400 slices of 1 A, 0.5 A pixels, {n_atoms} atoms "
                             f"(water slab + five 50k-atom clouds), one image",
                 "images_per_s": round(1 / t, 3), "ms_per_image": round(t * 1e3, 1),
                 "gpu_busy_frac": busy}
    progress(f"C3: {out['C3']['ms_per_image']} ms per image")
    return out


# ---------------------------------------------------------------------------
# CPU baselines
# ---------------------------------------------------------------------------
def cpu_baseline(mols, cfg, tem, budget_s: float = 12.0, max_images: int = 32):
    """CPU restatement of the reference's fftw backend on the host cores.

    oracle/fftw_port.c (C, OpenMP over all threads) runs whole images of the
    same workload until budget_s has elapsed; without it (no C compiler),
    the numpy oracle (single thread) runs one image.
    """
    sys.path.insert(0, HERE)
    from oracle import emsim_oracle as orc
    from emsim_amd import atoms as atm
    n, nS, dz, d = cfg["n"], cfg["slices"], cfg["dz"], cfg["pixel"]
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    psi0 = float(np.float32(np.sqrt(tem.electron_dose * d * d)))
    prepared = []
    for m in mols[:max_images]:
        mol, zs, counts = atm.sort_elements_and_count(atm.centralize(m), cfg["elements"])
        prepared.append((mol.coordinates, counts))
    sf = orc.scattering_factors_half(cfg["elements"], n, n, d, d)
    try:
        from oracle import port
        port.load()
    except Exception:
        port = None
    if port is not None:
        done, t0 = 0, time.perf_counter()
        while done < len(prepared) and (done == 0 or time.perf_counter() - t0 < budget_s):
            coords, counts = prepared[done]
            img = port.simulate(coords, counts, nS, n, dz, d, sf, lam, gamma, tem.cs_mm,
                                tem.defocus, tem.aperture, psi0)
            assert np.isfinite(img).all()
            done += 1
        dt = time.perf_counter() - t0
        return {"value": done / dt, "unit": "images/s", "cores": port.threads(), "kind": "port",
                "sample": f"{done} images of the workload ({n}x{n}, all {nS} slices) through "
                          f"oracle/fftw_port.c (C/OpenMP restatement of the fftw backend, own "
                          f"radix-2 FFT: FFTW is not installed), {dt:.1f} s"}
    coords, counts = prepared[0]
    t0 = time.perf_counter()
    hist = orc.bin_atoms_multi_slices(coords, counts, nS, n, n, dz, d, d)
    v = orc.make_multi_slices_fftw32(hist, sf).astype(np.float32)
    w = orc.multislice_fftw32(orc.init_wave(n, n, d, d, tem.electron_dose), v, dz, lam, gamma,
                              d, d)
    img = orc.intensity(orc.lens_propagate_fftw32(w, lam, tem.cs_mm, tem.defocus,
                                                  tem.aperture, d, d))
    dt = time.perf_counter() - t0
    assert np.isfinite(img).all()
    return {"value": 1.0 / dt, "unit": "images/s", "cores": 1, "kind": "port",
            "sample": f"1 image of the workload ({n}x{n}, all {nS} slices) through the numpy "
                      f"oracle (float32 fftw semantics, single-threaded pocketfft), {dt:.1f} s"}


def _numpy_image(orc, mol, elements, n, nS, dz, d, tem, sample_slices=None):
    """The reference's numpy backend on one molecule (slice_builder_numpy.py,
    wave_propagator_numpy.py): binning, make_multi_slices, multislice, lens,
    |psi|^2; with sample_slices, only that many slices are built and
    propagated (the caller extrapolates)."""
    from emsim_amd import atoms as atm
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    mol, zs, counts = atm.sort_elements_and_count(atm.centralize(mol), elements)
    ns = nS if sample_slices is None else sample_slices
    sf = orc.scattering_factors_half(list(zs), n, n, d, d)
    hist = orc.bin_atoms_multi_slices(mol.coordinates, counts, nS, n, n, dz, d, d)[:, :ns]
    v = orc.make_multi_slices_numpy(hist, sf)
    w = orc.multislice_numpy(orc.init_wave(n, n, d, d, tem.electron_dose), v, dz, lam, gamma,
                             d, d)
    return np.abs(orc.lens_propagate_numpy(w, lam, tem.cs_mm, tem.defocus, tem.aperture,
                                           d, d)) ** 2


def cpu_rows(tem, mols):
    """The CPU restatements on the other configurations, each on a bounded
    sample (seconds), stated per row."""
    sys.path.insert(0, HERE)
    from oracle import emsim_oracle as orc
    from emsim_amd import atoms as atm
    from emsim_amd import synth
    rows = []
    lam, gamma = orc.wavelength_and_gamma(tem.beam_energy_kev)
    try:
        from oracle import port
        port.load()
        cores = port.threads()
    except Exception:
        port, cores = None, None

    def timed(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    # C1: five atoms, one 512^2 slice, Scherzer optics (numpy backend one-slice path)
    from emsim_amd import physics
    five = atm.centralize(synth.five_atoms())
    d1 = 50.0 / 512
    cs = 1.3
    df, ap = physics.scherzer_condition(lam, cs)

    def c1():
        mol, zs, counts = atm.sort_elements_and_count(five)
        hist = orc.bin_atoms_one_slice(mol.coordinates, counts, 512, 512, d1, d1)
        sf = orc.scattering_factors_half(list(zs), 512, 512, d1, d1)
        v = orc.make_one_slice_fftw32(hist, sf)
        w = orc.slice_transmit_numpy(orc.init_wave(512, 512, d1, d1, 20.0), v, lam, gamma)
        return np.abs(orc.lens_propagate_numpy(w, lam, cs, df, ap, d1, d1)) ** 2
    t = timed(c1)
    rows.append({"config": "C1 five atoms, 512^2, one slice", "path": "numpy backend",
                 "images_per_s": round(1 / t, 3), "cores": 1,
                 "sample": "whole image, median of 3"})

    # C2: 50k-atom cloud, 256^2, 45 slices of 2 A
    c2mol = atm.centralize(synth.protein_cloud(50000, 0))
    ns2 = int(np.ceil(float(c2mol.space[0]) / 2.0))
    t = timed(lambda: _numpy_image(orc, c2mol, [1, 6, 7, 8, 16], 256, ns2, 2.0, 1.0, tem), 1)
    rows.append({"config": f"C2 50k atoms, 256^2, {ns2} slices of 2 A", "path": "numpy backend",
                 "images_per_s": round(1 / t, 3), "cores": 1,
                 "sample": "whole image, median of 1 after a warm-up"})
    if port is not None:
        mol, zs, counts = atm.sort_elements_and_count(c2mol, [1, 6, 7, 8, 16])
        sf2 = orc.scattering_factors_half([1, 6, 7, 8, 16], 256, 256, 1.0, 1.0)
        psi0 = float(np.float32(np.sqrt(20.0)))
        t = timed(lambda: port.simulate(mol.coordinates, counts, ns2, 256, 2.0, 1.0, sf2, lam,
                                        gamma, tem.cs_mm, tem.defocus, tem.aperture, psi0))
        rows.append({"config": f"C2 50k atoms, 256^2, {ns2} slices of 2 A",
                     "path": "fftw backend restatement (oracle/fftw_port.c)",
                     "images_per_s": round(1 / t, 3), "cores": cores,
                     "sample": "whole image, median of 3"})

    # M: the benchmark workload, numpy backend on 10 of the 100 slices
    m0 = mols[0]
    t = timed(lambda: _numpy_image(orc, m0, [1, 6, 7, 8, 16], 512, 100, 1.0, 1.0, tem,
                                   sample_slices=10), 1)
    rows.append({"config": "M 512^2 x 100 slices (benchmark molecule 0)", "path": "numpy backend",
                 "images_per_s": round(1 / (t * 10), 4), "cores": 1,
                 "sample": "10 of the 100 slices built and propagated, time x 10"})

    # C3: 2048^2, 0.5 A pixels: 8 slices of the 400-slice slab, time x 50
    if port is not None:
        cut = synth.water_slab(8.0, 1024.0, 1024.0, seed=1)
        cloud = synth.protein_cloud(50000, 2001)
        keep = np.abs(cloud.coordinates[:, 0]) < 3.9
        mol = atm.concatenate([cut, atm.AtomList(cloud.elements[keep], cloud.coordinates[keep])])
        mol, zs, counts = atm.sort_elements_and_count(atm.centralize(mol), [1, 6, 7, 8, 16])
        sf3 = orc.scattering_factors_half([1, 6, 7, 8, 16], 2048, 2048, 0.5, 0.5)
        psi0 = float(np.float32(np.sqrt(20.0 * 0.25)))
        t0 = time.perf_counter()
        port.simulate(mol.coordinates, counts, 8, 2048, 1.0, 0.5, sf3, lam, gamma, tem.cs_mm,
                      tem.defocus, tem.aperture, psi0)
        t = time.perf_counter() - t0
        rows.append({"config": "C3 2048^2, 0.5 A pixels, 400 slices of 1 A",
                     "path": "fftw backend restatement (oracle/fftw_port.c)",
                     "images_per_s": round(1 / (t * 50), 5), "cores": cores,
                     "sample": f"8 slices of the slab ({len(mol)} atoms), one run of "
                               f"{t:.1f} s, time x 50"})
    return rows


# ---------------------------------------------------------------------------
def progress(msg: str) -> None:
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def calibrate_copy(mib: int):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 4
    print(json.dumps({"calibration": "copy", "bytes_read": n * 4, "bytes_written": n * 4,
                      "GBs": round(2 * n * 4 / dt / 1e9, 1)}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              f"(images would shard across GPUs with no collective)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import torch
    import emsim_amd
    from emsim_amd.backend import hip_ext
    from emsim_amd.batch import BatchPipe

    cfg = dict(n=args.n, slices=args.slices, pixel=1.0, dz=1.0, batch=args.batch,
               elements=[1, 6, 7, 8, 16], n_elems=5)
    tem = emsim_amd.tem.TEM(electron_dose=20, beam_energy_kev=200, cs_mm=1.3, defocus=700,
                            aperture=10e-3)
    if args.config_only:
        print(json.dumps(config_lines(tem, (args.config_only,))), flush=True)
        return
    mols, sizes = workload(args.batch)
    progress(f"workload: {args.batch} molecules, mean {np.mean(sizes):.0f} atoms")
    if args.learned_only:
        print(json.dumps(learned_bench(args, cfg, tem, mols, None)), flush=True)
        return
    pipe = BatchPipe(tem, cfg["pixel"], cfg["dz"], cfg["n"], cfg["slices"], cfg["elements"])
    prep = pipe.prepare_many(mols)
    prepared = pipe.upload(prep)
    run = pipe.run_unfused if args.unfused else pipe.run
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        run(prepared)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    for _ in range(args.steps):
        images = run(prepared)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernels = []
    if not args.no_timing:
        # per-launch events cost a few microseconds per launch, so they time
        # a second pass of the same steps, not the throughput above
        hip_ext.timing_collect()
        hip_ext.timing_enable(True)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            run(prepared)
        torch.cuda.synchronize()
        elapsed_instr = time.perf_counter() - t1
        hip_ext.timing_enable(False)
        kernels = hip_ext.timing_collect()

    assert torch.isfinite(images).all().item()
    n_images = args.batch * args.steps
    value = n_images / elapsed
    ms_per_step = elapsed / args.steps * 1e3
    progress(f"plain path: {value:.1f} images/s ({ms_per_step:.2f} ms per batch)")

    # atoms inside the simulated box (binned), for the spectrum pass's bytes
    half = np.array([cfg["slices"] * cfg["dz"], cfg["n"] * cfg["pixel"],
                     cfg["n"] * cfg["pixel"]]) / 2
    binned = sum(int(np.count_nonzero(np.all(np.abs(p.coords) < half, axis=1))) for p in prep)
    ctx = dict(cfg, n_atoms=int(sum(p.coords.shape[0] for p in prep)), binned_atoms=binned,
               live=live_columns(cfg["n"], cfg["pixel"]))
    traffic, traffic_note = _profile(args.traffic)
    valu, valu_note = _profile(args.valu)

    roofline = None
    ktable = []
    gpu_total = sum(k["total_ms"] for k in kernels)
    step_measured = 0.0
    step_measured_ok = traffic is not None and bool(kernels)
    for k in sorted(kernels, key=lambda r: -r["total_ms"]):
        b = compulsory_bytes(k["name"], ctx)
        mean_s = k["mean_ms"] * 1e-3
        row = {"kernel": k["name"], "launches_per_step": k["count"] / args.steps,
               "mean_us": round(k["mean_ms"] * 1e3, 2),
               "share": round(k["total_ms"] / gpu_total, 4),
               "compulsory_MB": None if not math.isfinite(b) else round(b / 1e6, 2),
               "compulsory_GBs": None if not math.isfinite(b) else round(b / mean_s / 1e9, 1)}
        t = _lookup(traffic, k["name"]) if traffic else None
        if t is not None:
            row["measured_MB"] = round(t["bytes"] / 1e6, 2)
            row["measured_GBs"] = round(t["bytes"] / mean_s / 1e9, 1)
            step_measured += t["bytes"] * k["count"] / args.steps
        elif traffic is not None:
            step_measured_ok = False
        v = _lookup(valu, k["name"]) if valu else None
        if v is not None and "SQ_INSTS_VALU" in v:
            from emsim_amd.profiling import VALU_ISSUE_PEAK
            row["valu_issue_frac"] = round(v["SQ_INSTS_VALU"] / (mean_s * VALU_ISSUE_PEAK), 3)
        ktable.append(row)
    if ktable:
        top = ktable[0]
        if top["compulsory_GBs"] is not None:
            roofline = {
                "bound": "hbm", "achieved": top["compulsory_GBs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(top["compulsory_GBs"] / HBM_PEAK_GBS, 4),
                "traffic": None if "measured_MB" not in top else round(top["measured_MB"] * 1e6),
                "kernel": top["kernel"], "mean_us": top["mean_us"],
                "compulsory_bytes": round(top["compulsory_MB"] * 1e6),
                "measured_frac": None if "measured_GBs" not in top
                else round(top["measured_GBs"] / HBM_PEAK_GBS, 4),
                "valu_issue_frac": top.get("valu_issue_frac"),
                "profiles": {"traffic": traffic_note or os.path.relpath(args.traffic, HERE),
                             "valu": valu_note or os.path.relpath(args.valu, HERE)},
                "limiter": "FP32 VALU issue (FFT butterflies) and LDS-exchange / barrier "
                           "latency; HBM traffic is a small fraction of peak",
            }
    step_hbm = None
    if kernels:
        comp = sum(r["compulsory_MB"] * 1e6 * r["launches_per_step"] for r in ktable
                   if r["compulsory_MB"] is not None)
        step_hbm = {
            "compulsory_bytes_per_step": round(comp),
            "compulsory_frac": round(comp / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "measured_bytes_per_step": round(step_measured) if step_measured_ok else None,
            "measured_frac": round(step_measured / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            if step_measured_ok else None,
            "dense_reference_model": {
                "bytes_per_image": round(dense_reference_bytes_per_image(cfg["n"], cfg["slices"],
                                                                         cfg["n_elems"])),
                "note": "the reference's dense passes (histograms, per-element spectra, "
                        "potentials in HBM); the fused path never moves these bytes, so "
                        "this is a model of the replaced algorithm, not a roofline"},
        }

    mean_atoms = float(np.mean(sizes))
    learned = None
    if not args.no_learned and os.path.exists(args.learned) and args.n == 512 and \
            args.slices == 100 and not args.unfused:
        del images
        torch.cuda.empty_cache()
        learned = learned_bench(args, cfg, tem, mols, value)
        progress(f"learned line: {learned['images_per_s']} images/s")

    e2e = None
    if not args.no_e2e and not args.unfused:
        e2e = end_to_end(pipe, mols, args.batch)
        progress(f"host-to-host line: {e2e['images_per_s']} images/s")

    configs = None
    if not args.no_configs and not args.unfused:
        torch.cuda.empty_cache()
        configs = config_lines(tem)
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(mols, cfg, tem)
        progress(f"CPU baseline: {cpu['value']:.3f} images/s")
    rows = None
    if not args.no_cpu_rows and args.n == 512 and args.slices == 100:
        rows = cpu_rows(tem, mols)
        progress("CPU rows done")

    line = {
        "metric": "images/s at 512^2 x 100 slices (plain multislice, full path atoms->image)",
        "value": round(value, 2),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/c64",
        "data": "synthetic",
        "config": {"workload": f"plain multislice {cfg['n']}^2 x {cfg['slices']} slices, "
                               f"dz 1 A, pixel 1 A, batch of {args.batch} independent "
                               f"molecules of {int(sizes.min())}-{int(sizes.max())} atoms "
                               f"(mean {mean_atoms:.0f}; H/C/N/O/S, seeds 20001-"
                               f"{20000 + args.batch}), random orientations, 200 keV, "
                               f"Cs 1.3 mm, df 700 A, 10 mrad, no water",
                   "batch": args.batch, "n": cfg["n"], "slices": cfg["slices"],
                   "mean_atoms": round(mean_atoms, 1)},
        "slices_per_s": round(value * cfg["slices"], 1),
        "gpu_busy_frac": round(gpu_total * 1e-3 / elapsed_instr, 4) if kernels else None,
        "roofline": roofline,
        "step_hbm": step_hbm,
        "kernels": ktable,
        "learned": learned,
        "end_to_end": e2e,
        "configs": configs,
        "cpu_baseline": cpu,
        "cpu_rows": rows,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
