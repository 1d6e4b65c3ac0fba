"""The C ABI boundary, checked without a GPU: the library loads, exports every
function include/emsim_hip.h declares, the ctypes binding declares exactly
that set, and calls that fail argument checks return status codes and
messages before touching a device."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "emsim_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(emsim_hip_\w+)\s*\(", text)))


def test_header_declares_the_expected_entry_points():
    names = declared_functions()
    for must in ("emsim_hip_slice_builder_create", "emsim_hip_bin_atoms", "emsim_hip_make_slices",
                 "emsim_hip_wave_propagator_create", "emsim_hip_multislice_propagate",
                 "emsim_hip_lens_propagate", "emsim_hip_simulate"):
        assert must in names


def test_abi_version_agrees_across_header_binding_and_library():
    """build() and the binding's import check both compare against
    hip_ext.ABI_VERSION; it must be the header's EMSIM_HIP_ABI_VERSION."""
    m = re.search(r"#define\s+EMSIM_HIP_ABI_VERSION\s+(\d+)", HEADER.read_text())
    assert m
    hip_ext = pytest.importorskip("emsim_amd.backend.hip_ext")
    assert hip_ext.ABI_VERSION == int(m.group(1)) == hip_ext.abi_version()


def test_library_exports_every_declared_symbol():
    hip_ext = pytest.importorskip("emsim_amd.backend.hip_ext")
    lib = ctypes.CDLL(str(hip_ext.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hip_ext.EXPORTED) == declared_functions()


def test_argument_errors_without_a_device():
    hip_ext = pytest.importorskip("emsim_amd.backend.hip_ext")
    h = ctypes.c_void_p()
    rc = hip_ext.raw("emsim_hip_wave_propagator_create")(ctypes.byref(h), 3000, 100, 1.0, 1.0,
                                                         0.025, 1.39)
    assert rc == hip_ext.EUNSUPPORTED and not h.value
    assert "sizes from 8 to 2048" in hip_ext.last_error()
    rc = hip_ext.raw("emsim_hip_wave_propagator_create")(ctypes.byref(h), 64, 64, -1.0, 1.0,
                                                         0.025, 1.39)
    assert rc == hip_ext.EINVAL
    with pytest.raises(ValueError, match="powers of two up to 4096"):
        hip_ext.call("emsim_hip_wave_propagator_create", ctypes.byref(h), 4, 64, 1.0, 1.0,
                     0.025, 1.39)
    rc = hip_ext.raw("emsim_hip_slice_builder_create")(ctypes.byref(h), None, 1, 1, 64, 64, 1.0,
                                                       1.0, 1.0, 0)
    assert rc == hip_ext.EINVAL
    assert hip_ext.raw("emsim_hip_intensity")(None, None, 10, None) == hip_ext.EINVAL
    assert hip_ext.fft_size_supported(512) and hip_ext.fft_size_supported(500)
    assert not hip_ext.fft_size_supported(5000) and not hip_ext.fft_size_supported(7)
    assert hip_ext.raw("emsim_hip_fft_size_supported")(500) == 2
    assert hip_ext.abi_version() == 2
    # learned correction: shapes are checked before any device allocation
    f = ctypes.c_void_p()
    assert hip_ext.raw("emsim_hip_fno_create")(ctypes.byref(f), 512, 15, 24, 16, 4) == \
        hip_ext.EUNSUPPORTED and not f.value
    assert hip_ext.raw("emsim_hip_fno_create")(ctypes.byref(f), 512, 15, 32, 200, 4) == \
        hip_ext.EINVAL
    assert hip_ext.raw("emsim_hip_fno_create")(ctypes.byref(f), 500, 15, 32, 16, 4) == \
        hip_ext.EUNSUPPORTED
    d = hip_ext.GemmDesc(m=4, n=4, k=4, batch=1)        # null operands
    assert hip_ext.raw("emsim_hip_gemm")(ctypes.addressof(d), None) == hip_ext.EINVAL
    d = hip_ext.GemmDesc(m=4, n=4, k=4, batch=1, a=1, b=1, c=1, c_type=hip_ext.DT_BF16,
                         split_k=2)                    # split-k partials must be float32
    assert hip_ext.raw("emsim_hip_gemm")(ctypes.addressof(d), None) == hip_ext.EINVAL
    assert hip_ext.raw("emsim_hip_fno_param_count")(None) == -1


def test_backend_registry():
    import emsim_amd
    emsim_amd.config.set_backend("hip")
    assert emsim_amd.config.get_current_backend().backend_type is emsim_amd.config.BackendType.HIP
    for name in ("numpy", "fftw", "cuda"):
        with pytest.raises(ValueError, match="reference emsim"):
            emsim_amd.config.set_backend(name)
    with pytest.raises(ValueError, match="unknown backend"):
        emsim_amd.config.set_backend("opencl")
    with pytest.raises(ValueError):
        emsim_amd.config.set_backend(3.0)
    emsim_amd.config.set_backend("hip")
