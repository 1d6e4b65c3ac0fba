"""Build the gfx950 HIP library ``emsim_amd/backend/hip_ext/libemsim_hip.so``.

Usage: python emsim_amd/csrc/build.py [--verbose]

hipcc cross-compiles for gfx950 without a GPU present.  Packed-FP32 VALU
instructions are disabled for all device code.
"""
import os
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
OUT_DIR = PKG / "backend" / "hip_ext"
LIB = OUT_DIR / "libemsim_hip.so"
SOURCES = [HERE / "emsim_hip.hip", HERE / "fno.hip"]
DEPS = SOURCES + sorted(HERE.glob("*.hpp")) + [PKG.parent / "include" / "emsim_hip.h"]

ARCH = os.environ.get("EMSIM_HIP_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, *map(str, SOURCES), "-o", str(tmp)]
    t0 = time.time()
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    if verbose:
        noise = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if noise:
            print("\n".join(noise))
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    build(verbose="--verbose" in sys.argv or "-v" in sys.argv)
