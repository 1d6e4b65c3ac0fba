"""Learned coarse-slice correction on the HIP backend.

The reference has no learned component; this module adds one beside the
plain path and keeps the plain path as its fallback and as the measure of its
error.

``FNO`` wraps the library's Fourier neural operator (include/emsim_hip.h,
``emsim_hip_fno_*``): 4 spectral layers of 32 channels keeping 16 modes per
axis by default, bf16 MFMA with float32 accumulation, forward, backward and
Adam all in hand-written HIP kernels.  Its input is the exit wave of a
coarse-slice multislice run (the shipped network: 4 slices of 25 A for a
100 A thick box) plus that run's projected potentials; its output is the
correction towards the fine-slice (1 A) exit wave:
psi = psi_coarse + psi0 * (out_r + i out_i).

``CoarseFinePairs`` makes training pairs with the plain fused GPU path (the
same atoms simulated at both slice thicknesses), ``train`` fits the network
on them, and ``LearnedPipe`` runs batches through coarse multislice + the
network, checking every batch (norm conservation of each corrected wave, and
one sampled image re-simulated with plain fine-slice multislice) and falling
back to plain fine-slice multislice for the whole batch when a check fails.
"""
import time
from ctypes import byref, c_int64, c_void_p
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np
import torch

from .backend import hip_ext
from .backend._device import device, ptr, stream_handle
from .batch import BatchPipe, PreparedBatch, PreparedMolecule
from .tem import TEM


def _bf16_to_f32(u16: np.ndarray) -> np.ndarray:
    return (u16.astype(np.uint32) << 16).view(np.float32)


class FNO:
    """Fourier neural operator on the GPU (parameters live in the library)."""

    def __init__(self, n: int, in_channels: int, channels: int = 32, modes: int = 16,
                 layers: int = 4):
        h = c_void_p()
        hip_ext.call("emsim_hip_fno_create", byref(h), int(n), int(in_channels), int(channels),
                     int(modes), int(layers))
        self.h = h
        self.n, self.cin, self.C, self.modes, self.L = n, in_channels, channels, modes, layers
        self.cin_pad = (in_channels + 15) // 16 * 16
        self.n_params = int(hip_ext.raw("emsim_hip_fno_param_count")(h))
        offs = (c_int64 * (4 + 3 * layers))()
        hip_ext.call("emsim_hip_fno_param_offsets", h, offs, 4 + 3 * layers)
        self.offsets = [int(o) for o in offs]
        self._ws = {}
        self._loss_ws = None
        self.step = 0

    def __del__(self):
        try:
            if getattr(self, "h", None):
                hip_ext.raw("emsim_hip_fno_destroy")(self.h)
                self.h = None
        except Exception:   # interpreter shutdown
            pass

    # -- parameters ---------------------------------------------------------
    def _blocks(self):
        C, m, cp = self.C, self.modes, self.cin_pad
        shapes = [("lift_w", (cp, C)), ("lift_b", (C,))]
        for l in range(self.L):
            shapes += [(("w", l), (C, C)), (("b", l), (C,)), (("r", l), (2 * m, m, C, C, 2))]
        shapes += [("proj_w", (C, 2)), ("proj_b", (2,))]
        return list(zip(self.offsets, shapes))

    def unflatten(self, flat: np.ndarray) -> dict:
        p = {"w": [None] * self.L, "b": [None] * self.L, "r": [None] * self.L}
        for off, (key, shape) in self._blocks():
            a = np.asarray(flat[off:off + int(np.prod(shape))], dtype=np.float64).reshape(shape)
            if isinstance(key, tuple):
                name, l = key
                p[name][l] = a[..., 0] + 1j * a[..., 1] if name == "r" else a
            else:
                p[key] = a
        return p

    def flatten(self, p: dict) -> np.ndarray:
        flat = np.zeros(self.n_params, dtype=np.float32)
        for off, (key, shape) in self._blocks():
            if isinstance(key, tuple):
                name, l = key
                a = p[name][l]
                if name == "r":
                    a = np.stack([np.real(a), np.imag(a)], axis=-1)
            else:
                a = p[key]
            a = np.asarray(a, dtype=np.float32)
            if a.shape != shape:
                raise ValueError(f"parameter {key}: shape {a.shape} != {shape}")
            flat[off:off + a.size] = a.reshape(-1)
        return flat

    def init_params(self, seed: int, proj_scale: float = 0.1) -> dict:
        """Uniform(+-1/sqrt(fan_in)) dense weights, spectral weights
        (1/C^2) U(0, 1) (re and im) as in Li et al.'s reference code, and a
        projection scaled by proj_scale so the initial correction is small."""
        rng = np.random.default_rng(seed)
        C, m = self.C, self.modes

        def uni(shape, fan_in):
            s = 1.0 / np.sqrt(fan_in)
            return rng.uniform(-s, s, size=shape)

        lift_w = np.zeros((self.cin_pad, C))
        lift_w[:self.cin] = uni((self.cin, C), self.cin)
        p = {"lift_w": lift_w, "lift_b": uni((C,), self.cin), "w": [], "b": [], "r": []}
        for _ in range(self.L):
            p["w"].append(uni((C, C), C))
            p["b"].append(uni((C,), C))
            p["r"].append((rng.uniform(0, 1, (2 * m, m, C, C)) +
                           1j * rng.uniform(0, 1, (2 * m, m, C, C))) / (C * C))
        p["proj_w"] = proj_scale * uni((C, 2), C)
        p["proj_b"] = np.zeros(2)
        return p

    def set_params(self, p) -> None:
        flat = p if isinstance(p, (np.ndarray, torch.Tensor)) else self.flatten(p)
        t = torch.as_tensor(flat, dtype=torch.float32).to(device()).contiguous()
        if t.numel() != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {t.numel()}")
        hip_ext.call("emsim_hip_fno_set_params", self.h, ptr(t), stream_handle())
        torch.cuda.current_stream().synchronize()

    def params_flat(self) -> np.ndarray:
        t = torch.empty(self.n_params, dtype=torch.float32, device=device())
        hip_ext.call("emsim_hip_fno_get_params", self.h, ptr(t), stream_handle())
        return t.cpu().numpy()

    def get_params(self) -> dict:
        return self.unflatten(self.params_flat())

    def grads(self) -> dict:
        t = torch.empty(self.n_params, dtype=torch.float32, device=device())
        hip_ext.call("emsim_hip_fno_get_grads", self.h, ptr(t), stream_handle())
        return self.unflatten(t.cpu().numpy())

    def save(self, path: str, coarse_dz: float = None) -> None:
        """Parameters and shape; coarse_dz (the slice thickness of the coarse
        simulation the network was trained on) is stored with them."""
        extra = {} if coarse_dz is None else {"coarse_dz": np.array([float(coarse_dz)])}
        np.savez(path, params=self.params_flat(),
                 config=np.array([self.n, self.cin, self.C, self.modes, self.L, self.step]),
                 **extra)

    @classmethod
    def load(cls, path: str) -> "FNO":
        """A saved network; ``coarse_dz`` is None for files saved without it
        (networks trained on 8 A coarse slices)."""
        with np.load(path, allow_pickle=False) as z:
            n, cin, C, m, L, step = (int(v) for v in z["config"])
            f = cls(n, cin, C, m, L)
            f.set_params(z["params"])
            f.step = step
            f.coarse_dz = float(z["coarse_dz"][0]) if "coarse_dz" in z.files else None
        return f

    # -- device calls --------------------------------------------------------
    def workspace(self, batch: int, training: bool) -> torch.Tensor:
        key = (batch, bool(training))
        t = self._ws.get(key)
        if t is None:
            nbytes = hip_ext.raw("emsim_hip_fno_workspace_size")(self.h, batch, 1 if training else 0)
            t = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device())
            self._ws = {key: t}   # keep one workspace
        return t

    def pack(self, wave: torch.Tensor, slices: torch.Tensor, psi0: float, sigma: float,
             out: torch.Tensor = None) -> torch.Tensor:
        B, n = wave.shape[0], self.n
        self._check(wave, (B, n, n), torch.complex64, "wave")
        ns = self.cin - 2
        self._check(slices, (B, ns, n, n), torch.float32, "slices")
        if out is None:
            out = torch.empty((B, n, n, self.cin_pad), dtype=torch.bfloat16, device=device())
        self._check(out, (B, n, n, self.cin_pad), torch.bfloat16, "out")
        hip_ext.call("emsim_hip_fno_pack_input", self.h, ptr(wave), ptr(slices), ns, float(psi0),
                     float(sigma), B, ptr(out), stream_handle())
        return out

    def forward(self, x: torch.Tensor, training: bool = False,
                out: torch.Tensor = None) -> torch.Tensor:
        B = x.shape[0]
        self._check(x, (B, self.n, self.n, self.cin_pad), torch.bfloat16, "x")
        if out is None:
            out = torch.empty((B, self.n, self.n, 2), dtype=torch.float32, device=device())
        self._check(out, (B, self.n, self.n, 2), torch.float32, "out")
        ws = self.workspace(B, training)
        hip_ext.call("emsim_hip_fno_forward", self.h, ptr(x), ptr(out), B, 1 if training else 0,
                     ptr(ws), ws.numel(), stream_handle())
        return out

    def _check(self, t: torch.Tensor, shape, dtype, name: str) -> None:
        if t.dtype != dtype or tuple(t.shape) != tuple(shape) or not t.is_contiguous() or \
                not t.is_cuda:
            raise ValueError(f"{name} must be a contiguous {dtype} GPU tensor of shape "
                             f"{tuple(shape)}, got {t.dtype} {tuple(t.shape)}")

    def infer(self, wave: torch.Tensor, slices: torch.Tensor, psi0: float, sigma: float,
              out: torch.Tensor = None) -> torch.Tensor:
        """forward(pack(wave, slices, ...)) with the packing fused into the
        lifting (same values); falls back to the two steps for network shapes
        the fused kernels do not serve."""
        B, n = wave.shape[0], self.n
        self._check(wave, (B, n, n), torch.complex64, "wave")
        self._check(slices, (B, self.cin - 2, n, n), torch.float32, "slices")
        if out is None:
            out = torch.empty((B, n, n, 2), dtype=torch.float32, device=device())
        self._check(out, (B, n, n, 2), torch.float32, "out")
        ws = self.workspace(B, False)
        rc = hip_ext.raw("emsim_hip_fno_infer")(self.h, ptr(wave), ptr(slices), self.cin - 2,
                                               float(psi0), float(sigma), ptr(out), B, ptr(ws),
                                               ws.numel(), stream_handle())
        if rc == hip_ext.EUNSUPPORTED:
            return self.forward(self.pack(wave, slices, psi0, sigma), out=out)
        hip_ext.check(rc, "emsim_hip_fno_infer")
        return out

    def backward(self, x: torch.Tensor, dout: torch.Tensor) -> None:
        B = x.shape[0]
        self._check(x, (B, self.n, self.n, self.cin_pad), torch.bfloat16, "x")
        self._check(dout, (B, self.n, self.n, 2), torch.float32, "dout")
        ws = self.workspace(B, True)
        hip_ext.call("emsim_hip_fno_backward", self.h, ptr(x), ptr(dout), B, ptr(ws),
                     ws.numel(), stream_handle())

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
                  weight_decay: float = 0.0) -> None:
        self.step += 1
        hip_ext.call("emsim_hip_fno_adam", self.h, float(lr), float(beta1), float(beta2),
                     float(eps), float(weight_decay), self.step, stream_handle())

    def apply(self, coarse: torch.Tensor, out: torch.Tensor, psi0: float,
              wave: torch.Tensor = None) -> torch.Tensor:
        B = coarse.shape[0]
        self._check(coarse, (B, self.n, self.n), torch.complex64, "coarse")
        self._check(out, (B, self.n, self.n, 2), torch.float32, "out")
        if wave is None:
            wave = torch.empty_like(coarse)
        self._check(wave, (B, self.n, self.n), torch.complex64, "wave")
        hip_ext.call("emsim_hip_fno_apply", self.h, ptr(coarse), ptr(out), float(psi0), ptr(wave),
                     coarse.shape[0], stream_handle())
        return wave

    def loss(self, pred: torch.Tensor, target: torch.Tensor, psi0: float, grad: bool = True):
        return wave_loss(pred, target, psi0, grad)


def wave_loss(pred: torch.Tensor, target: torch.Tensor, psi0: float = 1.0, grad: bool = False):
    """(loss (B, 3) = [|pred - target|^2 / |target|^2, |target|^2, |pred|^2],
    dout or None), computed by the library's reduction kernels."""
    B = pred.shape[0]
    npix = pred[0].numel()
    for t, name in ((pred, "pred"), (target, "target")):
        if t.dtype != torch.complex64 or tuple(t.shape) != tuple(pred.shape) or \
                not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"{name} must be a contiguous complex64 GPU tensor shaped like pred")
    loss = torch.empty((B, 3), dtype=torch.float32, device=device())
    dout = torch.empty((B,) + tuple(pred.shape[1:]) + (2,), dtype=torch.float32,
                       device=device()) if grad else None
    nbytes = hip_ext.raw("emsim_hip_fno_loss_workspace_size")(npix, B)
    ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device())
    hip_ext.call("emsim_hip_fno_loss", ptr(pred), ptr(target), npix, B, float(psi0), ptr(loss),
                 None if dout is None else ptr(dout), ptr(ws), ws.numel(), stream_handle())
    return loss, dout


# ---------------------------------------------------------------------------
# training pairs and training
# ---------------------------------------------------------------------------
def subset(batch: PreparedBatch, idx: Sequence[int]) -> PreparedBatch:
    """The molecules idx of a prepared batch (atoms stay on the GPU)."""
    parts, offs = [], [0]
    for i in idx:
        lo, hi = int(batch.mol_off[i]), int(batch.mol_off[i + 1])
        parts.append(batch.coords[lo:hi])
        offs.append(offs[-1] + hi - lo)
    coords = torch.cat(parts) if parts else batch.coords[:0]
    return PreparedBatch(coords.contiguous(), np.asarray(offs, dtype=np.int64),
                         np.ascontiguousarray(batch.counts[list(idx)]),
                         [batch.labels[i] for i in idx])


class CoarseFinePairs:
    """The same molecules through plain fused multislice at two slice
    thicknesses: (coarse exit wave, coarse potentials) -> fine exit wave."""

    def __init__(self, microscope: TEM, n: int = 512, pixel: float = 1.0,
                 fine_dz: float = 1.0, fine_slices: int = 100, coarse_dz: float = 8.0,
                 coarse_slices: int = 13, elements: Sequence[int] = (1, 6, 7, 8, 16)):
        self.fine = BatchPipe(microscope, pixel, fine_dz, n, fine_slices, elements)
        self.coarse = BatchPipe(microscope, pixel, coarse_dz, n, coarse_slices, elements)
        self.n, self.coarse_slices = n, coarse_slices
        prop = self.fine.propagator
        self.psi0 = float(prop.init_wave_value(microscope.electron_dose))
        self.sigma = float(np.float32(prop.wave_length) * np.float32(prop.relativity_gamma))

    def make(self, batch: PreparedBatch, fine: bool = True):
        cw, cs = self.coarse.exit_waves_and_slices(batch)
        fw = self.fine.exit_waves(batch) if fine else None
        return cw, cs, fw


def train(fno: FNO, pairs: CoarseFinePairs, batches: Callable[[int], List[PreparedMolecule]],
          steps: int, lr: float = 2e-3, lr_final: float = 1e-4, log_every: int = 50,
          log: Callable[[str], None] = print) -> List[dict]:
    """Adam on the mean relative squared exit-wave error; batches(step) gives
    the molecules of each step.  Cosine learning-rate decay lr -> lr_final."""
    history = []
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        pb = pairs.fine.upload(batches(step))
        cw, cs, fw = pairs.make(pb)
        x = fno.pack(cw, cs, pairs.psi0, pairs.sigma)
        out = fno.forward(x, training=True)
        pred = fno.apply(cw, out, pairs.psi0)
        loss, dout = fno.loss(pred, fw, pairs.psi0, grad=True)
        fno.backward(x, dout)
        frac = (step - 1) / max(steps - 1, 1)
        cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
        fno.adam_step(cur)
        if step % log_every == 0 or step == steps:
            base, _ = wave_loss(cw, fw)
            l = loss[:, 0].sqrt().mean().item()
            b = base[:, 0].sqrt().mean().item()
            history.append({"step": step, "rel_l2": l, "coarse_rel_l2": b, "lr": cur,
                            "elapsed_s": time.perf_counter() - t0})
            log(f"step {step}: rel L2 {l:.5f} (coarse alone {b:.5f}), lr {cur:.2e}, "
                f"{time.perf_counter() - t0:.0f} s")
    return history


# ---------------------------------------------------------------------------
# gated learned pipeline
# ---------------------------------------------------------------------------
@dataclass
class GateReport:
    fallback: bool = False
    norm_dev: List[float] = field(default_factory=list)   # | |psi|^2 / |psi_coarse|^2 - 1 |
    sampled: int = -1                                       # image re-simulated finely
    sampled_rel_l2: float = float("nan")                    # learned vs plain fine, that image
    reason: str = ""


class LearnedPipe:
    """Coarse multislice + FNO correction, with plain fine-slice multislice as
    the fallback (taken when a norm-conservation or sampled-image residual
    check fails) and as the measure of the error."""

    def __init__(self, fno: FNO, pairs: CoarseFinePairs, norm_tol: float = 0.01,
                 check_tol: float = 0.01, check_every: int = 1):
        if fno.n != pairs.n or fno.cin != 2 + pairs.coarse_slices:
            raise ValueError("network and pipelines disagree on size or slice count")
        if pairs.fine.add_water or pairs.coarse.add_water:
            raise ValueError("the correction network is trained on molecules without water")
        self.fno, self.pairs = fno, pairs
        self.norm_tol, self.check_tol, self.check_every = norm_tol, check_tol, check_every
        self._calls = 0
        # The sampled image's plain fine-slice run uses its own pipe (own
        # buffers) on a second stream of default priority, concurrent with the
        # batch's coarse run and network; a single image leaves most of the
        # GPU idle, so this hides most of its cost.
        f = pairs.fine
        self._check_pipe = BatchPipe(f.microscope, f.pixel_size, f.dz, f.n, f.n_slices,
                                     f.elements, f.symmetric_bandlimit)
        self._side = torch.cuda.Stream(device=device())

    def exit_waves(self, batch: PreparedBatch, check: bool = True):
        p = self.pairs
        main = torch.cuda.current_stream()
        sampled = None
        if check and self._calls % self.check_every == 0:
            s = self._calls // self.check_every % len(batch)
            one = subset(batch, [s])
            self._side.wait_stream(main)
            with torch.cuda.stream(self._side):
                fine_s = self._check_pipe.exit_waves(one)
            one.coords.record_stream(self._side)
            sampled = (s, fine_s)
        cw, cs, _ = p.make(batch, fine=False)
        out = self.fno.infer(cw, cs, p.psi0, p.sigma)
        wave = self.fno.apply(cw, out, p.psi0)
        rep = GateReport()
        if not check:
            return wave, rep
        self._calls += 1
        stats, _ = wave_loss(wave, cw)
        dev = (stats[:, 2] / stats[:, 1] - 1).abs().cpu().numpy()
        rep.norm_dev = dev.tolist()
        if sampled is not None:
            s, fine_s = sampled
            main.wait_stream(self._side)
            fine_s.record_stream(main)
            err, _ = wave_loss(wave[s:s + 1], fine_s)
            rep.sampled, rep.sampled_rel_l2 = s, float(err[0, 0].sqrt().item())
        if (dev > self.norm_tol).any():
            rep.fallback, rep.reason = True, f"norm check: max deviation {dev.max():.3g}"
        elif rep.sampled >= 0 and rep.sampled_rel_l2 > self.check_tol:
            rep.fallback = True
            rep.reason = f"sampled image {rep.sampled}: rel L2 {rep.sampled_rel_l2:.3g} vs fine"
        if rep.fallback:
            wave = p.fine.exit_waves(batch)
        return wave, rep

    def run(self, batch: PreparedBatch, check: bool = True):
        """(B, n, n) float32 images and the gate report."""
        wave, rep = self.exit_waves(batch, check)
        mic = self.pairs.fine.microscope
        img = self.pairs.fine.propagator.lens_intensity(wave, mic.cs_mm, mic.defocus,
                                                        mic.aperture, overwrite_wave=True)
        return img, rep
