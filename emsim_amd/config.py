"""Backend registry (``emsim/config.py:4-69``).

Same shape as the reference: a ``BackendType`` enum, a ``Backend`` exposing
``one_slice_builder`` / ``multi_slices_builder`` / ``wave_propagator``
classes, and ``set_backend`` / ``get_current_backend``.  This package
implements the ``"hip"`` backend; the reference's numpy, fftw and cuda
backends stay in the reference package, so naming them here raises
ValueError like any other unknown name.  The backend modules are imported
on first use, so ``import emsim_amd`` works on a machine without a GPU.
"""
from enum import Enum


class BackendType(Enum):
    NUMPY = 1
    CUDA = 2
    FFTW = 3
    HIP = 4


_PROVIDED_ELSEWHERE = {BackendType.NUMPY: "numpy", BackendType.CUDA: "cuda",
                       BackendType.FFTW: "fftw"}


class Backend(object):
    def __init__(self, backend_type: BackendType):
        if backend_type is not BackendType.HIP:
            name = _PROVIDED_ELSEWHERE.get(backend_type, str(backend_type))
            raise ValueError(f"emsim_amd: backend '{name}' is provided by the reference emsim "
                             f"package; emsim_amd implements 'hip'")
        self.backend_type = backend_type
        self._slice_builder_module = None
        self._wave_propagator_module = None

    def _load(self):
        if self._slice_builder_module is None:
            from .backend import slice_builder_hip, wave_propagator_hip
            self._slice_builder_module = slice_builder_hip
            self._wave_propagator_module = wave_propagator_hip

    @property
    def one_slice_builder(self):
        self._load()
        return self._slice_builder_module.OneSliceBuilder

    @property
    def multi_slices_builder(self):
        self._load()
        return self._slice_builder_module.MultiSlicesBuilder

    @property
    def wave_propagator(self):
        self._load()
        return self._wave_propagator_module.WavePropagator

    def __repr__(self):
        return f"<emsim backend {str(self.backend_type)}>"


_current_backend = Backend(BackendType.HIP)


def get_current_backend():
    return _current_backend


def set_backend(backend="hip"):
    global _current_backend
    type_dict = {"numpy": BackendType.NUMPY, "cuda": BackendType.CUDA,
                 "fftw": BackendType.FFTW, "hip": BackendType.HIP}
    if type(backend) is str:
        backend_type = type_dict.get(backend)
        if backend_type is None:
            raise ValueError(f"emsim: unknown backend: {backend}")
    elif type(backend) is BackendType:
        backend_type = backend
    else:
        raise ValueError("pass str or BackendType as argumen to set_backend")
    _current_backend = Backend(backend_type)
