"""hip slice builders: the MI355X counterpart of ``slice_builder_fftw.py`` /
``slice_builder_cuda.py`` in the reference backend directory.

Same classes, constructors and methods as the reference's builders
(``emsim/backend/slice_builder_base.py:8-163``); every method runs HIP
kernels through ``hip_ext`` and returns float32 torch tensors on the GPU.
Binning reproduces numpy.histogramdd on the reference's float64 edges
bit-exactly; the potential build fuses the reference's R2C -> form-factor
sum -> C2R and its numpy band-limit + clip into one R2C, one column pass and
one C2R per plane.
"""
import logging
import threading
from collections import OrderedDict
from ctypes import POINTER, byref, c_int64, c_void_p
from typing import List

import numpy as np
import torch

from . import hip_ext
from ..physics import water_num_dens
from ._device import as_device, device, ptr, stream_handle, workspace
from .slice_builder_base import MultiSlicesBuilderBase, OneSliceBuilderBase

logger = logging.getLogger(__name__)

# Bytes of half-spectrum workspace one make_slices call may use; more slices
# than fit are transformed in chunks.
SLICES_WORKSPACE_BUDGET = 1 << 31


class _BuilderHandle:
    """Owns one emsim_hip_slice_builder (device SF table and edge tables)."""

    def __init__(self, sf: np.ndarray, n_slices, n1, n2, dz, d1, d2, one_slice):
        device()  # raise early without a GPU
        sf = np.ascontiguousarray(sf, dtype=np.float32)
        h = c_void_p()
        hip_ext.call("emsim_hip_slice_builder_create", byref(h), sf.ctypes.data_as(c_void_p),
                     sf.shape[0], int(n_slices), int(n1), int(n2), float(dz), float(d1),
                     float(d2), 1 if one_slice else 0)
        self.h = h
        self.n_elems = sf.shape[0]
        self.n_slices, self.n1, self.n2 = int(n_slices), int(n1), int(n2)
        self.one_slice = bool(one_slice)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and hip_ext is not None:
            try:
                hip_ext.raw("emsim_hip_slice_builder_destroy")(h)
            except Exception:       # interpreter shutdown
                pass
            self.h = None

    # -- operations --------------------------------------------------------
    def bin_atoms(self, coords, counts, out: torch.Tensor = None,
                  ws: torch.Tensor = None) -> torch.Tensor:
        coords_t = as_device(coords, torch.float32, "atom_coordinates")
        if coords_t.ndim != 2 or coords_t.shape[1] != 3:
            raise ValueError(f"atom coordinates must be (n_atoms, 3), got {tuple(coords_t.shape)}")
        counts_np = np.ascontiguousarray(
            counts.detach().cpu().numpy() if isinstance(counts, torch.Tensor) else np.asarray(counts),
            dtype=np.int64).reshape(-1)
        if counts_np.shape[0] != self.n_elems:
            raise ValueError(f"expected {self.n_elems} element counts, got {counts_np.shape[0]}")
        n_atoms = coords_t.shape[0]
        shape = ((self.n_elems, self.n1, self.n2) if self.one_slice
                 else (self.n_elems, self.n_slices, self.n1, self.n2))
        hist = _out_buffer(out, shape, coords_t.device)
        ws_bytes = hip_ext.raw("emsim_hip_bin_workspace_size")(self.h, n_atoms)
        ws = _workspace(ws, ws_bytes)
        hip_ext.call("emsim_hip_bin_atoms", self.h, ptr(coords_t),
                     counts_np.ctypes.data_as(POINTER(c_int64)), n_atoms, ptr(hist), ptr(ws),
                     ws_bytes, stream_handle())
        return hist

    def slices_workspace_bytes(self) -> int:
        per_slice = hip_ext.raw("emsim_hip_slices_workspace_size")(self.h, 1)
        chunk = max(1, min(self.n_slices, SLICES_WORKSPACE_BUDGET // max(per_slice, 1)))
        return hip_ext.raw("emsim_hip_slices_workspace_size")(self.h, chunk)

    def make_slices(self, hist, symmetric_bandlimit: bool, out: torch.Tensor = None,
                    ws: torch.Tensor = None) -> torch.Tensor:
        hist_t = as_device(hist, torch.float32, "atom_histograms")
        expect = ((self.n_elems, self.n1, self.n2) if self.one_slice
                  else (self.n_elems, self.n_slices, self.n1, self.n2))
        if tuple(hist_t.shape) != expect:
            raise ValueError(f"atom histograms must have shape {expect}, got {tuple(hist_t.shape)}")
        out_shape = (self.n1, self.n2) if self.one_slice else (self.n_slices, self.n1, self.n2)
        out = _out_buffer(out, out_shape, hist_t.device)
        ws_bytes = self.slices_workspace_bytes()
        ws = _workspace(ws, ws_bytes)
        hip_ext.call("emsim_hip_make_slices", self.h, ptr(hist_t), ptr(out),
                     1 if symmetric_bandlimit else 0, ptr(ws), ws.numel(), stream_handle())
        return out


def _out_buffer(out, shape, dev) -> torch.Tensor:
    if out is None:
        return torch.empty(shape, dtype=torch.float32, device=dev)
    if (tuple(out.shape) != tuple(shape) or out.dtype != torch.float32 or out.device != dev
            or not out.is_contiguous()):
        raise ValueError(f"out must be a contiguous float32 GPU tensor of shape {tuple(shape)}")
    return out


def _workspace(ws, nbytes) -> torch.Tensor:
    if ws is None:
        return workspace(nbytes)
    if ws.dtype != torch.uint8 or ws.numel() < nbytes or not ws.is_contiguous():
        raise ValueError(f"workspace must be a contiguous uint8 tensor of >= {nbytes} bytes")
    return ws


_HANDLES = OrderedDict()
_HANDLES_LOCK = threading.Lock()
_HANDLES_MAX = 32


def _handle(sf, key, *args) -> _BuilderHandle:
    with _HANDLES_LOCK:
        h = _HANDLES.get(key)
        if h is not None:
            _HANDLES.move_to_end(key)
            return h
    h = _BuilderHandle(sf, *args)
    with _HANDLES_LOCK:
        _HANDLES[key] = h
        while len(_HANDLES) > _HANDLES_MAX:
            _HANDLES.popitem(last=False)
    return h


class OneSliceBuilder(OneSliceBuilderBase):
    def __init__(self, unique_elements: List[int], n1: int, n2: int, d1: float, d2: float):
        logger.debug("using hip OneSliceBuilder")
        super().__init__(unique_elements, n1, n2, d1, d2)
        key = (tuple(int(z) for z in sorted(unique_elements)), 1, self.n1, self.n2, 0.0,
               self.d1, self.d2, True)
        self.backend = _handle(self.scattering_factors, key, 1, self.n1, self.n2, 0.0,
                               self.d1, self.d2, True)

    def bin_atoms_one_slice(self, atom_coordinates_sorted_by_elems, unique_elements_count):
        return self.backend.bin_atoms(atom_coordinates_sorted_by_elems, unique_elements_count)

    def make_one_slice(self, atom_histograms_one_slice, symmetric_bandlimit: bool = True):
        return self.backend.make_slices(atom_histograms_one_slice, symmetric_bandlimit)


class MultiSlicesBuilder(MultiSlicesBuilderBase):
    def __init__(self, unique_elements: List[int], n_slices: int, n1: int, n2: int,
                 dz: float, d1: float, d2: float):
        logger.debug("using hip MultiSlicesBuilder")
        super().__init__(unique_elements, n_slices, n1, n2, dz, d1, d2)
        key = (tuple(int(z) for z in sorted(unique_elements)), self.n_slice, self.n1, self.n2,
               self.dz, self.d1, self.d2, False)
        self.backend = _handle(self.scattering_factors, key, self.n_slice, self.n1, self.n2,
                               self.dz, self.d1, self.d2, False)

    def bin_atoms_multi_slices(self, atom_coordinates_sorted_by_elems, unique_elements_count):
        return self.backend.bin_atoms(atom_coordinates_sorted_by_elems, unique_elements_count)

    def add_water(self, atom_histograms, seed: int = None):
        """Poisson water in every voxel empty of all elements, in place.

        As ``slice_builder_base.py:134-147``: means water_num_dens d1 d2 dz
        (O) and twice that (H), hydrogen and oxygen must be among the
        builder's elements.  ``seed`` defaults to a draw from numpy's global
        RNG, so ``np.random.seed`` makes runs repeatable.
        """
        hist = as_device(atom_histograms, torch.float32, "atom_histograms")
        expect = (self.backend.n_elems, self.n_slice, self.n1, self.n2)
        if tuple(hist.shape) != expect:
            raise ValueError(f"atom histograms must have shape {expect}, got {tuple(hist.shape)}")
        zs = [int(z) for z in self.unique_elements]
        if 1 not in zs or 8 not in zs:
            raise ValueError("add_water needs H (Z=1) and O (Z=8) among the builder's elements")
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 63 - 1, dtype=np.int64))
        voxel_water = water_num_dens * self.d1 * self.d2 * self.dz
        hip_ext.call("emsim_hip_add_water", self.backend.h, ptr(hist), zs.index(1), zs.index(8),
                     float(voxel_water), int(seed) & ((1 << 64) - 1), stream_handle())
        return hist

    def make_multi_slices(self, atom_histograms, symmetric_bandlimit: bool = True):
        return self.backend.make_slices(atom_histograms, symmetric_bandlimit)
