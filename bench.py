"""Throughput benchmark of the hip multislice path on one MI355X.

Workload (the headline metric's configuration): 512 x 512 waves, 1 A pixel,
100 slices of 1 A, 200 keV, Cs 1.3 mm, defocus 700 A, aperture 10 mrad,
synthetic protein-like clouds (H/C/N/O/S, 1 A minimum spacing) of
log-uniform 10k-100k atoms in random orientations, no water.  One step is
one pass of the full path over one batch of molecules whose atoms are
already in GPU memory: binning -> potential slices -> multislice -> lens ->
|psi|^2 (emsim_amd.batch.BatchPipe).

Prints one JSON line.  ``value`` comes from the timed steps run without
instrumentation.  The same number of steps is then run again with every
launch bracketed by HIP events on its stream (the per-kernel table):
``roofline`` describes the kernel with the largest share of GPU time there,
its algorithmic bytes per launch (fixed per-pixel figures, see
kernel_bytes_per_launch) divided by its mean launch duration.
``cpu_baseline`` times the CPU oracle on a bounded sample of the same
workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--slices", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--unfused", action="store_true",
                    help="time the reference-shaped kernels (histogram, make_slices, multislice)")
    ap.add_argument("--no-timing", action="store_true",
                    help="do not bracket launches with HIP events")
    ap.add_argument("--traffic", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                      "profiles", "traffic_512x100_b64.json"),
                    help="rocprofv3 FETCH/WRITE summary for roofline.traffic")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 4 device-to-device copies of MIB MiB (a known-size "
                         "reference for HBM byte counters) and exit")
    ap.add_argument("--learned", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                      "emsim_amd", "data", "fno_512.npz"),
                    help="trained correction network for the learned coarse-slice line")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-only", action="store_true",
                    help="run only the learned coarse-slice line (prints its JSON object)")
    return ap.parse_args()


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
    """Coarse (8 A) fused multislice + FNO correction + lens, gated, on the
    same molecules; error against plain fine-slice multislice."""
    import torch
    from emsim_amd.backend import hip_ext
    from emsim_amd.learned import FNO, CoarseFinePairs, LearnedPipe, wave_loss

    fno = FNO.load(args.learned)
    coarse_dz = 8.0
    coarse_slices = fno.cin - 2
    pairs = CoarseFinePairs(tem, n=cfg["n"], pixel=cfg["pixel"], fine_dz=cfg["dz"],
                            fine_slices=cfg["slices"], coarse_dz=coarse_dz,
                            coarse_slices=coarse_slices, elements=cfg["elements"])
    lp = LearnedPipe(fno, pairs)
    batch = pairs.fine.upload([pairs.fine.prepare(m) for m in mols])
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
    gpu_total = sum(k["total_ms"] for k in kernels)
    top = sorted(kernels, key=lambda r: -r["total_ms"])[:8]
    return {
        "images_per_s": round(value, 2),
        "speedup_vs_plain": round(value / plain_ips, 3) if plain_ips else None,
        "images_per_s_fine_check_every_4th_batch": round(value4, 2),
        "speedup_vs_plain_fine_check_every_4th_batch":
            round(value4 / plain_ips, 3) if plain_ips else None,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "slices_per_s_fine_equivalent": round(value * cfg["slices"], 1),
        "coarse_dz": coarse_dz, "coarse_slices": coarse_slices,
        "network": {"layers": fno.L, "channels": fno.C, "modes": fno.modes, "dtype": "bf16 MFMA, f32 accumulate",
                    "train_steps": fno.step},
        "exit_wave_rel_l2_vs_plain": {"mean": float(err.mean()), "max": float(err.max()),
                                      "coarse_alone_mean": float(base.mean()),
                                      "coarse_alone_max": float(base.max())},
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
        "gpu_busy_frac": round(gpu_total * 1e-3 / elapsed_instr, 4) if kernels else None,
        "kernels": [{"kernel": k["name"], "launches": k["count"],
                     "mean_us": round(k["mean_ms"] * 1e3, 2),
                     "ms_per_step": round(k["total_ms"] / args.steps, 3)} for k in top],
    }


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


def workload(batch: int, n_atoms_range=(10000, 100000), n_base: int = 8, seed: int = 10000):
    """Batch of molecules: n_base base clouds, each image a random rotation of one."""
    from emsim_amd import synth
    from emsim_amd.utils.rot import random_uniform_quaternions
    rng = np.random.default_rng(seed)
    lo, hi = np.log(n_atoms_range[0]), np.log(n_atoms_range[1])
    sizes = np.exp(rng.uniform(lo, hi, n_base)).astype(int)
    bases = [synth.protein_cloud(int(s), seed + 1 + i) for i, s in enumerate(sizes)]
    quats = random_uniform_quaternions(batch, rng=np.random.default_rng(2))
    return [synth.rotated(bases[i % n_base], quats[i]) for i in range(batch)], sizes


def kernel_bytes_per_launch(name: str, n: int, cfg) -> float:
    """Algorithmic HBM bytes of one launch of a kernel in this workload.

    Per real pixel per slice (h = (n/2+1)/n, nE elements), the reference's
    passes are charged at:
      binning              histogram out 4 nE
      R2C                  histogram in 4 nE + half spectra out 8h nE
      spectrum column pass half spectra in 8h nE + out 8h
      C2R + clip           half spectrum in 8h + V out 4
      pass A               wave in 8 + V 4 + wave out 8 = 20
      pass B               wave in 8 + out 8 = 16
      lens epilogue        wave in 8 + image out 4 = 12
    The fused kernels are charged for the passes they replace:
      k_spec_sparse        binning + R2C + spectrum pass = nE (8 + 16h) + 8h
      k_slice_pass_a       C2R + pass A = 8h + 24
    Fused kernels and pass A/B cover one slice of the whole batch per launch;
    the unfused potential-build kernels cover all slices of one molecule.
    """
    B, npix, nS, nE = cfg["batch"], cfg["n"] ** 2, cfg["slices"], cfg["n_elems"]
    h = (cfg["n"] // 2 + 1) / cfg["n"]
    per = {
        "k_spec_sparse": (nE * (8 + 16 * h) + 8 * h) * npix * B,
        "k_slice_pass_a": (8 * h + 24.0) * npix * B,
        "k_row_transmit_fft": 20.0 * npix * B,
        "k_col_fft_mul_ifft": 16.0 * npix * B,
        "k_row_fft": 16.0 * npix * B,
        "k_row_ifft_intensity": 12.0 * npix * B,
        "k_row_r2c_pairs": (4.0 * nE + 8 * h * nE) * npix * nS,
        "k_col_spectrum": (8 * h * nE + 8 * h) * npix * nS + 4 * h * nE * npix,
        "k_row_c2r_pairs": (8 * h + 4.0) * npix * nS,
        "k_bin_tiles": 4.0 * nE * npix * nS,
    }
    for k, v in per.items():
        if name.startswith(k):
            return v
    return float("nan")


def measured_traffic(kernel: str, path: str):
    """HBM bytes per launch of `kernel` from a rocprofv3 FETCH/WRITE summary
    (emsim_amd.profiling), used only if it was measured on this exact library."""
    try:
        from emsim_amd.profiling import library_sha256
        data = json.loads(open(path).read())
        if data["library_sha256"] != library_sha256():
            return None
        base = kernel.split("<")[0]
        for name, rec in data["kernels"].items():
            if name.split("<")[0].split("(")[0].strip() == base:
                return round(rec["bytes"])
    except Exception:
        return None
    return None


def cpu_baseline(mols, cfg, tem, budget_s: float = 12.0, max_images: int = 32):
    """CPU restatement of the reference's fftw backend on the host cores.

    oracle/fftw_port.c (C, OpenMP over all threads) runs whole images of the
    same workload until budget_s has elapsed; without it (no C compiler),
    the numpy oracle (single thread) runs one image.
    """
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
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
    mols, sizes = workload(args.batch)
    if args.learned_only:
        print(json.dumps(learned_bench(args, cfg, tem, mols, None)), flush=True)
        return
    pipe = BatchPipe(tem, cfg["pixel"], cfg["dz"], cfg["n"], cfg["slices"], cfg["elements"])
    prepared = pipe.upload([pipe.prepare(m) for m in mols])
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

    roofline = None
    ktable = []
    gpu_total = sum(k["total_ms"] for k in kernels)
    for k in sorted(kernels, key=lambda r: -r["total_ms"]):
        b = kernel_bytes_per_launch(k["name"], k["n"], cfg)
        gbs = b / (k["mean_ms"] * 1e-3) / 1e9 if math.isfinite(b) else None
        ktable.append({"kernel": f"{k['name']}", "launches": k["count"],
                       "mean_us": round(k["mean_ms"] * 1e3, 2),
                       "share": round(k["total_ms"] / gpu_total, 4),
                       "alg_GBs": None if gbs is None else round(gbs, 1)})
    if ktable:
        top = next((r for r in ktable if r["alg_GBs"] is not None), None)
        if top is not None:
            roofline = {"bound": "hbm", "achieved": top["alg_GBs"], "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(top["alg_GBs"] / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(top["kernel"], args.traffic),
                        "kernel": top["kernel"], "mean_us": top["mean_us"]}

    learned = None
    if not args.no_learned and os.path.exists(args.learned) and args.n == 512 and \
            args.slices == 100 and not args.unfused:
        del images
        pipe = None
        torch.cuda.empty_cache()
        learned = learned_bench(args, cfg, tem, mols, value)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(mols, cfg, tem)

    line = {
        "metric": "images/s at 512^2 x 100 slices (plain multislice, full path atoms->image)",
        "value": round(value, 2),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/c64",
        "data": "synthetic",
        "config": {"workload": f"plain multislice {cfg['n']}^2 x {cfg['slices']} slices, "
                               f"dz 1 A, pixel 1 A, batch {args.batch} molecules of "
                               f"{int(sizes.min())}-{int(sizes.max())} atoms (H/C/N/O/S), "
                               f"random orientations, 200 keV, Cs 1.3 mm, df 700 A, 10 mrad",
                   "batch": args.batch, "n": cfg["n"], "slices": cfg["slices"]},
        "slices_per_s": round(value * cfg["slices"], 1),
        "gpu_busy_frac": round(gpu_total * 1e-3 / elapsed_instr, 4) if kernels else None,
        "roofline": roofline,
        "kernels": ktable,
        "learned": learned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
