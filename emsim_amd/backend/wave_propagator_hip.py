"""hip wave propagator: the MI355X counterpart of ``wave_propagator_fftw.py`` /
``wave_propagator_cuda.py`` in the reference backend directory.

Same constructor and six methods as the reference's propagators
(``emsim/backend/wave_propagator_base.py:10-44``), each returning a new
complex64 torch tensor on the GPU.  Waves may carry a leading batch axis:
(n1, n2) or (B, n1, n2); a batch runs through every kernel in one launch.
``lens_intensity`` is an addition that fuses the lens with |psi|^2.
"""
import logging
import threading
from collections import OrderedDict
from ctypes import byref, c_void_p

import numpy as np
import torch

from . import hip_ext
from ._device import as_device, device, ptr, stream_handle
from .wave_propagator_base import WavePropagatorBase

logger = logging.getLogger(__name__)


class _PropagatorHandle:
    def __init__(self, n1, n2, d1, d2, lam, gamma):
        device()
        h = c_void_p()
        hip_ext.call("emsim_hip_wave_propagator_create", byref(h), int(n1), int(n2),
                     float(np.float32(d1)), float(np.float32(d2)), float(np.float32(lam)),
                     float(np.float32(gamma)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and hip_ext is not None:
            try:
                hip_ext.raw("emsim_hip_wave_propagator_destroy")(h)
            except Exception:       # interpreter shutdown
                pass
            self.h = None


_HANDLES = OrderedDict()
_HANDLES_LOCK = threading.Lock()


def _handle(*key) -> _PropagatorHandle:
    with _HANDLES_LOCK:
        h = _HANDLES.get(key)
        if h is not None:
            _HANDLES.move_to_end(key)
            return h
    h = _PropagatorHandle(*key)
    with _HANDLES_LOCK:
        _HANDLES[key] = h
        while len(_HANDLES) > 32:
            _HANDLES.popitem(last=False)
    return h


class WavePropagator(WavePropagatorBase):
    def __init__(self, n1: int, n2: int, d1: float, d2: float, beam_energy_kev: float):
        logger.debug("using hip WavePropagator")
        super().__init__(n1, n2, d1, d2, beam_energy_kev)
        self.backend = _handle(self.wave_shape[0], self.wave_shape[1], self.d1, self.d2,
                               self.wave_length, self.relativity_gamma)

    @property
    def n1(self):
        return self.wave_shape[0]

    @property
    def n2(self):
        return self.wave_shape[1]

    # -- argument handling -------------------------------------------------
    def _wave(self, wave, name="wave"):
        w = as_device(wave, torch.complex64, name)
        if w.ndim == 2:
            batched = False
            w3 = w.unsqueeze(0)
        elif w.ndim == 3:
            batched = True
            w3 = w
        else:
            raise ValueError(f"{name} must be (n1, n2) or (batch, n1, n2), got {tuple(w.shape)}")
        if tuple(w3.shape[1:]) != self.wave_shape:
            raise ValueError(f"{name} has lateral shape {tuple(w3.shape[1:])}, "
                             f"propagator expects {self.wave_shape}")
        return w3, batched

    def _out(self, w3, batched):
        return w3 if batched else w3[0]

    def _slices(self, slices, batch: int, name: str):
        """-> (tensor (..., nS, n1, n2) float32, n_slices, per-image stride)."""
        s = as_device(slices, torch.float32, name)
        npix = self.wave_shape[0] * self.wave_shape[1]
        if s.ndim == 3 and tuple(s.shape[1:]) == self.wave_shape:
            return s, s.shape[0], 0
        if s.ndim == 4 and tuple(s.shape[2:]) == self.wave_shape and s.shape[0] == batch:
            return s, s.shape[1], s.shape[1] * npix
        raise ValueError(f"{name} must be (n_slices, n1, n2) or (batch, n_slices, n1, n2) with "
                         f"lateral shape {self.wave_shape}, got {tuple(s.shape)}")

    # -- reference API -----------------------------------------------------
    def init_wave(self, electron_dose: float, batch: int = None):
        shape = self.wave_shape if batch is None else (int(batch),) + self.wave_shape
        out = torch.empty(shape, dtype=torch.complex64, device=device())
        hip_ext.call("emsim_hip_init_wave", self.backend.h,
                     float(self.init_wave_value(electron_dose)), ptr(out),
                     1 if batch is None else int(batch), stream_handle())
        return out

    def slice_transmit(self, wave, aslice):
        w3, batched = self._wave(wave)
        s = as_device(aslice, torch.float32, "aslice")
        npix = self.wave_shape[0] * self.wave_shape[1]
        if s.ndim == 2 and tuple(s.shape) == self.wave_shape:
            stride = 0
        elif s.ndim == 3 and s.shape[0] == w3.shape[0] and tuple(s.shape[1:]) == self.wave_shape:
            stride = npix
        else:
            raise ValueError(f"aslice must be (n1, n2) or (batch, n1, n2), got {tuple(s.shape)}")
        out = torch.empty_like(w3)
        hip_ext.call("emsim_hip_slice_transmit", self.backend.h, ptr(w3), ptr(s), stride,
                     ptr(out), w3.shape[0], stream_handle())
        return self._out(out, batched)

    def space_propagate(self, wave, dz: float):
        w3, batched = self._wave(wave)
        out = torch.empty_like(w3)
        hip_ext.call("emsim_hip_space_propagate", self.backend.h, ptr(w3), float(dz), ptr(out),
                     w3.shape[0], stream_handle())
        return self._out(out, batched)

    def singleslice_propagate(self, wave, aslice, dz: float):
        w3, batched = self._wave(wave)
        s = as_device(aslice, torch.float32, "aslice")
        if s.ndim == 2:
            s = s.unsqueeze(0)
        elif s.ndim == 3 and batched and s.shape[0] == w3.shape[0]:
            s = s.unsqueeze(1)
        return self._multislice(w3, batched, s, float(dz))

    def multislice_propagate(self, wave, slices, dz: float):
        w3, batched = self._wave(wave)
        return self._multislice(w3, batched, slices, float(dz))

    def _multislice(self, w3, batched, slices, dz):
        s, n_slices, stride = self._slices(slices, w3.shape[0], "slices")
        out = torch.empty_like(w3)
        hip_ext.call("emsim_hip_multislice_propagate", self.backend.h, ptr(w3), ptr(s), n_slices,
                     stride, dz, ptr(out), w3.shape[0], stream_handle())
        return self._out(out, batched)

    def lens_propagate(self, wave, cs_mm: float, defocus: float, aperture: float):
        w3, batched = self._wave(wave)
        out = torch.empty_like(w3)
        hip_ext.call("emsim_hip_lens_propagate", self.backend.h, ptr(w3), float(cs_mm),
                     float(defocus), float(aperture), ptr(out), w3.shape[0], stream_handle())
        return self._out(out, batched)

    # -- additions ---------------------------------------------------------
    def lens_intensity(self, wave, cs_mm: float, defocus: float, aperture: float,
                       overwrite_wave: bool = False):
        """|lens_propagate(wave)|^2 as float32, without materialising the image wave."""
        w3, batched = self._wave(wave)
        scratch = w3 if overwrite_wave else torch.empty_like(w3)
        img = torch.empty(w3.shape, dtype=torch.float32, device=w3.device)
        hip_ext.call("emsim_hip_lens_intensity", self.backend.h, ptr(w3), float(cs_mm),
                     float(defocus), float(aperture), ptr(scratch), ptr(img), w3.shape[0],
                     stream_handle())
        return img if batched else img[0]


def intensity(wave) -> torch.Tensor:
    """|psi|^2 of a complex64 device tensor (``emsim/pipe.py:76,82``)."""
    w = as_device(wave, torch.complex64, "wave")
    out = torch.empty(w.shape, dtype=torch.float32, device=w.device)
    hip_ext.call("emsim_hip_intensity", ptr(w), ptr(out), w.numel(), stream_handle())
    return out
