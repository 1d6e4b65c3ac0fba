"""ctypes binding of ``libemsim_hip.so`` (C ABI: ``include/emsim_hip.h``).

This is the thin layer between the Python backend classes and the HIP
library: it loads the shared object that sits next to this file, declares
every exported function's signature, and turns non-zero return codes into
Python exceptions carrying the library's message.  Device buffers are torch
tensors; only their data pointers and the current HIP stream cross the
boundary.  ctypes releases the GIL for the duration of each call.

Importing this module raises ImportError when the library has not been
built, as the reference's extension wrappers do
(``emsim/backend/slice_builder_fftw.py:7-10``).
"""
import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int64, c_size_t, c_uint64,
                    c_void_p)
from pathlib import Path

# EMSIM_HIP_LIB may name another build of the same library (e.g. one compiled
# with different tile constants) to load instead of the in-tree one.
LIB_PATH = Path(os.environ.get("EMSIM_HIP_LIB") or
                Path(__file__).resolve().parent / "libemsim_hip.so").resolve()

if not LIB_PATH.exists():
    raise ImportError(f"hip extension {LIB_PATH} cannot be found. Compile it first "
                      f"(python emsim_amd/csrc/build.py)")

_lib = ctypes.CDLL(str(LIB_PATH))

# the ABI the signatures below describe (EMSIM_HIP_ABI_VERSION in the header)
ABI_VERSION = 2
if _lib.emsim_hip_abi_version() != ABI_VERSION:
    raise ImportError(f"{LIB_PATH} reports ABI {_lib.emsim_hip_abi_version()}, this binding "
                      f"expects {ABI_VERSION}: rebuild it (python emsim_amd/csrc/build.py)")

EINVAL, EUNSUPPORTED, EWORKSPACE, EDEVICE, ENOMEM = -1, -2, -3, -4, -5

_SIGNATURES = {
    "emsim_hip_abi_version": (c_int, []),
    "emsim_hip_last_error": (c_char_p, []),
    "emsim_hip_fft_size_supported": (c_int, [c_int]),
    "emsim_hip_bin_edges": (c_int, [c_int, c_double, POINTER(c_double), POINTER(c_float), c_int,
                                    POINTER(c_int)]),
    "emsim_hip_slice_builder_create": (c_int, [POINTER(c_void_p), c_void_p, c_int, c_int, c_int,
                                               c_int, c_double, c_double, c_double, c_int]),
    "emsim_hip_slice_builder_destroy": (c_int, [c_void_p]),
    "emsim_hip_bin_workspace_size": (c_size_t, [c_void_p, c_int64]),
    "emsim_hip_bin_atoms": (c_int, [c_void_p, c_void_p, POINTER(c_int64), c_int64, c_void_p,
                                    c_void_p, c_size_t, c_void_p]),
    "emsim_hip_add_water": (c_int, [c_void_p, c_void_p, c_int, c_int, c_double, c_uint64,
                                    c_void_p]),
    "emsim_hip_slices_workspace_size": (c_size_t, [c_void_p, c_int]),
    "emsim_hip_make_slices": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_size_t,
                                      c_void_p]),
    "emsim_hip_wave_propagator_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_float,
                                                 c_float, c_float, c_float]),
    "emsim_hip_wave_propagator_destroy": (c_int, [c_void_p]),
    "emsim_hip_init_wave": (c_int, [c_void_p, c_float, c_void_p, c_int, c_void_p]),
    "emsim_hip_slice_transmit": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int,
                                         c_void_p]),
    "emsim_hip_space_propagate": (c_int, [c_void_p, c_void_p, c_float, c_void_p, c_int,
                                          c_void_p]),
    "emsim_hip_multislice_propagate": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int64,
                                               c_float, c_void_p, c_int, c_void_p]),
    "emsim_hip_lens_propagate": (c_int, [c_void_p, c_void_p, c_float, c_float, c_float,
                                         c_void_p, c_int, c_void_p]),
    "emsim_hip_lens_intensity": (c_int, [c_void_p, c_void_p, c_float, c_float, c_float,
                                         c_void_p, c_void_p, c_int, c_void_p]),
    "emsim_hip_intensity": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "emsim_hip_simulate_workspace_size": (c_size_t, [c_void_p, c_int, c_int64, c_int]),
    "emsim_hip_simulate": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, POINTER(c_int64),
                                   POINTER(c_int64), c_int, c_float, c_float, c_float, c_float,
                                   c_int, c_void_p, c_void_p, c_int, c_void_p, c_size_t,
                                   c_void_p]),
    "emsim_hip_simulate_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                       c_int, c_float, c_float, c_float, c_float, c_int,
                                       c_void_p, c_void_p, c_int, c_void_p, c_size_t, c_void_p]),
    "emsim_hip_simulate_water_workspace_size": (c_size_t, [c_void_p, c_int, c_int64, c_int]),
    "emsim_hip_simulate_water": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, POINTER(c_int64),
                                         POINTER(c_int64), c_int, c_float, c_float, c_float,
                                         c_float, c_int, c_int, c_int, c_double,
                                         POINTER(c_uint64), c_void_p, c_void_p, c_int, c_void_p,
                                         c_size_t, c_void_p]),
    "emsim_hip_simulate_raw_workspace_size": (c_size_t, [c_void_p, c_int, c_int64, c_int]),
    "emsim_hip_simulate_raw": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                       POINTER(c_int64), c_int, c_float, c_float, c_float,
                                       c_float, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                       c_size_t, c_void_p]),
    "emsim_hip_gemm": (c_int, [c_void_p, c_void_p]),
    "emsim_hip_fno_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int]),
    "emsim_hip_fno_destroy": (None, [c_void_p]),
    "emsim_hip_fno_param_count": (c_int64, [c_void_p]),
    "emsim_hip_fno_param_offsets": (c_int, [c_void_p, POINTER(c_int64), c_int]),
    "emsim_hip_fno_set_params": (c_int, [c_void_p, c_void_p, c_void_p]),
    "emsim_hip_fno_get_params": (c_int, [c_void_p, c_void_p, c_void_p]),
    "emsim_hip_fno_get_grads": (c_int, [c_void_p, c_void_p, c_void_p]),
    "emsim_hip_fno_workspace_size": (c_size_t, [c_void_p, c_int, c_int]),
    "emsim_hip_fno_pack_input": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_float, c_float,
                                         c_int, c_void_p, c_void_p]),
    "emsim_hip_fno_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                      c_size_t, c_void_p]),
    "emsim_hip_fno_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_size_t,
                                       c_void_p]),
    "emsim_hip_fno_adam": (c_int, [c_void_p, c_float, c_float, c_float, c_float, c_float, c_int,
                                   c_void_p]),
    "emsim_hip_fno_force_generic": (c_int, [c_int]),
    "emsim_hip_fno_infer": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_float, c_float,
                                    c_void_p, c_int, c_void_p, c_size_t, c_void_p]),
    "emsim_hip_fno_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_int,
                                    c_void_p]),
    "emsim_hip_fno_loss_workspace_size": (c_size_t, [c_int64, c_int]),
    "emsim_hip_fno_loss": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_float, c_void_p,
                                   c_void_p, c_void_p, c_size_t, c_void_p]),
    "emsim_hip_timing_enable": (c_int, [c_int]),
    "emsim_hip_orient_workspace_size": (c_size_t, [c_int]),
    "emsim_hip_orient_batch": (c_int, [c_void_p, c_int64, c_void_p, c_float, c_float, c_float,
                                       c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "emsim_hip_orient_batch_noalloc": (c_int, [c_void_p, c_int64, c_void_p, c_float, c_float,
                                               c_float, c_int, c_void_p, c_void_p]),
    "emsim_hip_tuning_set": (c_int, [c_char_p, c_int]),
    "emsim_hip_tuning_get": (c_int, [c_char_p]),
    "emsim_hip_timing_collect": (c_int, [c_void_p, c_size_t, POINTER(c_size_t)]),
}

EXPORTED = tuple(_SIGNATURES)

DT_F32, DT_BF16 = 0, 1


class GemmDesc(ctypes.Structure):
    """``emsim_gemm_desc``: one batched strided GEMM (see include/emsim_hip.h)."""
    _fields_ = [
        ("m", c_int64), ("n", c_int64), ("k", c_int64), ("batch", c_int64),
        ("a", c_void_p), ("a_bs", c_int64), ("a_sm", c_int64), ("a_sk", c_int64),
        ("a_type", ctypes.c_int32),
        ("b", c_void_p), ("b_bs", c_int64), ("b_sk", c_int64), ("b_sn", c_int64),
        ("b_type", ctypes.c_int32),
        ("k2", c_int64),
        ("a2", c_void_p), ("a2_bs", c_int64), ("a2_sm", c_int64), ("a2_sk", c_int64),
        ("a2_type", ctypes.c_int32),
        ("b2", c_void_p), ("b2_bs", c_int64), ("b2_sk", c_int64), ("b2_sn", c_int64),
        ("b2_type", ctypes.c_int32),
        ("c", c_void_p), ("c_bs", c_int64), ("c_sm", c_int64), ("c_sn", c_int64),
        ("c_type", ctypes.c_int32),
        ("bias", c_void_p),
        ("act", ctypes.c_int32),
        ("z", c_void_p), ("z_type", ctypes.c_int32),
        ("beta", c_float),
        ("split_k", c_int64), ("c_split_stride", c_int64),
    ]

for _name, (_res, _args) in _SIGNATURES.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


class HipExtError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def last_error() -> str:
    msg = _lib.emsim_hip_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    """Map a library status to the exception the reference API would raise."""
    if rc == 0:
        return
    msg = f"{what}: {last_error()}"
    if rc in (EINVAL, EUNSUPPORTED, EWORKSPACE):
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise HipExtError(rc, msg)


def call(name: str, *args) -> int:
    """Invoke an exported function by name and check its status."""
    check(getattr(_lib, name)(*args), name)
    return 0


def raw(name: str):
    return getattr(_lib, name)


def abi_version() -> int:
    return _lib.emsim_hip_abi_version()


def fft_size_supported(n: int) -> bool:
    return bool(_lib.emsim_hip_fft_size_supported(int(n)))


def bin_edges(n: int, d: float):
    """(float64 edges, float32 round-up edges) exactly as the library builds them."""
    import numpy as np
    length = c_int(0)
    check(_lib.emsim_hip_bin_edges(int(n), float(d), None, None, 0, ctypes.byref(length)),
          "bin_edges")
    e64 = np.empty(length.value, dtype=np.float64)
    e32 = np.empty(length.value, dtype=np.float32)
    check(_lib.emsim_hip_bin_edges(int(n), float(d),
                                   e64.ctypes.data_as(POINTER(c_double)),
                                   e32.ctypes.data_as(POINTER(c_float)),
                                   length.value, ctypes.byref(length)), "bin_edges")
    return e64, e32


def timing_enable(on: bool = True) -> None:
    """Bracket every kernel launch with HIP events on its stream."""
    check(_lib.emsim_hip_timing_enable(1 if on else 0), "timing_enable")


def timing_collect() -> list:
    """Per-(kernel, length) launch statistics recorded since the last collect.

    Returns dicts with name, n, count, total_ms, min_ms, max_ms, mean_ms.
    """
    need = c_size_t(0)
    cap = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(cap)
        check(_lib.emsim_hip_timing_collect(buf, cap, ctypes.byref(need)), "timing_collect")
        if need.value <= cap:
            break
        cap = need.value   # pragma: no cover - more than 64 KiB of distinct kernels
    rows = []
    for line in buf.value.decode().splitlines():
        name, n, count, total, lo, hi = line.split("\t")
        name = name.strip("()")
        rows.append(dict(name=name, n=int(n), count=int(count), total_ms=float(total),
                         min_ms=float(lo), max_ms=float(hi), mean_ms=float(total) / int(count)))
    return rows
