"""Write tests/golden/c3_slab_2048x400.npz: the large-slab configuration
through the CPU oracle (oracle/fftw_port.c, the fftw backend's float32
algorithm), all 400 slices.

Workload: 2048 x 2048 wave, 0.5 A pixels, 400 slices of 1 A, 200 keV,
Cs 1.3 mm, defocus 700 A, aperture 10 mrad, dose 20 e/A^2; atoms =
emsim_amd.synth.large_slab(seed=1) (water slab 400 x 1024 x 1024 A plus
five 50k-atom protein-like clouds), centred, elements H/C/N/O/S.

The full image (16 MB) is too large to keep, so the fixture holds what the
GPU test compares: a 256 x 256 crop of the image and a 128 x 128 crop of
the exit wave around the centre, the 64 x 64 sums of 32 x 32 image blocks
(float64), and the image's total and sum of squares (float64).

    python tests/golden/make_c3_fixture.py        # a few minutes on 8 cores
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from emsim_amd import atoms as atm  # noqa: E402
from emsim_amd import synth  # noqa: E402
from oracle import emsim_oracle as orc  # noqa: E402
from oracle import port  # noqa: E402

N, NS, PIXEL, DZ = 2048, 400, 0.5, 1.0
CS, DEFOCUS, APERTURE, DOSE = 1.3, 700.0, 10e-3, 20.0
ELEMS = [1, 6, 7, 8, 16]
IMG_CROP = (slice(896, 1152), slice(896, 1152))
WAVE_CROP = (slice(960, 1088), slice(960, 1088))
BLOCK = 32


def summarize(image, wave):
    """The compared statistics of a full image and exit wave."""
    img64 = image.astype(np.float64)
    blocks = img64.reshape(N // BLOCK, BLOCK, N // BLOCK, BLOCK).sum(axis=(1, 3))
    return {"image_crop": np.ascontiguousarray(image[IMG_CROP], dtype=np.float32),
            "wave_crop": np.ascontiguousarray(wave[WAVE_CROP], dtype=np.complex64),
            "block_sums": blocks, "image_sum": np.array([img64.sum()]),
            "image_sumsq": np.array([(img64 * img64).sum()])}


def main():
    t0 = time.time()
    mol = atm.centralize(synth.large_slab(seed=1))
    mol, _, counts = atm.sort_elements_and_count(mol, ELEMS)
    print(f"{len(mol)} atoms in {time.time() - t0:.0f} s", flush=True)
    lam, gamma = orc.wavelength_and_gamma(200.0)
    sf = orc.scattering_factors_half(ELEMS, N, N, PIXEL, PIXEL)
    psi0 = float(np.float32(np.sqrt(DOSE * PIXEL * PIXEL)))
    t0 = time.time()
    image, wave = port.simulate(mol.coordinates, counts, NS, N, DZ, PIXEL, sf, lam, gamma, CS,
                                DEFOCUS, APERTURE, psi0, exit_wave=True)
    print(f"oracle in {time.time() - t0:.0f} s on {port.threads()} threads", flush=True)
    out = Path(__file__).resolve().parent / "c3_slab_2048x400.npz"
    np.savez_compressed(out, n_atoms=np.array([len(mol)]), counts=np.asarray(counts),
                        **summarize(image, wave))
    print(f"wrote {out} ({out.stat().st_size / 1e3:.0f} kB)")


if __name__ == "__main__":
    main()
