"""ctypes front of oracle/libfftw_port.so -- TEST INFRASTRUCTURE ONLY.

The C/OpenMP restatement of the reference's fftw backend (fftw_port.c), used
as the multi-core CPU baseline of bench.py and checked against
emsim_oracle.py by tests/test_oracle_port.py.
"""
import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libfftw_port.so"


def load():
    if not LIB.exists():
        subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
    lib = ctypes.CDLL(str(LIB))
    lib.port_simulate.argtypes = [c_void_p, POINTER(c_int64), c_int, c_int, c_int, c_double,
                                  c_double, c_void_p, c_float, c_float, c_float, c_float, c_float,
                                  c_float, c_int, c_void_p]
    lib.port_simulate.restype = c_int
    lib.port_simulate_ex.argtypes = lib.port_simulate.argtypes + [c_void_p]
    lib.port_simulate_ex.restype = c_int
    lib.port_threads.restype = c_int
    return lib


def simulate(coords, counts, n_slices, n, dz, d, sf, lam, gamma, cs_mm, defocus, aperture,
             psi0, bandlimit=True, exit_wave=False):
    """One image through the whole path; returns (n, n) float32, and with
    exit_wave also the (n, n) complex64 wave before the lens."""
    lib = load()
    coords = np.ascontiguousarray(coords, dtype=np.float32)
    counts = np.ascontiguousarray(counts, dtype=np.int64)
    sf = np.ascontiguousarray(sf, dtype=np.float32)
    out = np.empty((n, n), dtype=np.float32)
    wave = np.empty((n, n), dtype=np.complex64) if exit_wave else None
    rc = lib.port_simulate_ex(coords.ctypes.data_as(c_void_p),
                              counts.ctypes.data_as(POINTER(c_int64)), len(counts), int(n_slices),
                              int(n), float(dz), float(d), sf.ctypes.data_as(c_void_p), float(lam),
                              float(gamma), float(cs_mm), float(defocus), float(aperture),
                              float(psi0), 1 if bandlimit else 0, out.ctypes.data_as(c_void_p),
                              None if wave is None else wave.ctypes.data_as(c_void_p))
    if rc != 0:
        raise ValueError("port_simulate: bin edges are not n + 1 (non-dyadic voxel size)")
    return (out, wave) if exit_wave else out


def threads() -> int:
    return load().port_threads()
