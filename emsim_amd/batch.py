"""Device-resident batched image pipeline.

The reference simulates one molecule at a time (``emsim/simulator.py:71-91``
calls ``emsim/pipe.py:79-84`` per molecule, and ``Pipe`` rebuilds its
propagator three times per image, ``pipe.py:31-34``).  ``BatchPipe`` runs
the same per-image algorithm -- centralize, bin, build slices, multislice,
lens, |psi|^2 -- for a batch of molecules whose atoms are already in GPU
memory.  Unlike ``Pipe`` it passes ``n_slices`` to the slice builder, so
every image has the same slice count.

``run`` / ``exit_waves`` use the fused kernels (``emsim_hip_simulate``):
atoms are bucketed once per batch, and each slice costs three launches for
the whole batch (spectrum from the atom buckets, potential C2R fused into the
transmission row pass, propagator column pass), with the lens folded into the
last slice.  ``run_unfused`` is the same computation through the
reference-shaped steps (dense histogram, ``make_multi_slices``,
``multislice_propagate``, ``lens_intensity``), kept as a cross-check.
"""
import os
from ctypes import POINTER, c_int64
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import atoms as atm
from .backend import hip_ext
from .backend._device import as_device, device, ptr, stream_handle
from .backend.slice_builder_hip import MultiSlicesBuilder
from .backend.wave_propagator_hip import WavePropagator
from .tem import TEM


def pool_map(fn, items, workers: int = None) -> list:
    """[fn(x) for x in items] on a thread pool (numpy releases the GIL in
    the sorts, reductions and copies of the host preparation); at most 16
    threads, the CPU share a GPU process is sized for."""
    items = list(items)
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    workers = max(1, min(int(workers), len(items)))
    if workers == 1:
        return [fn(x) for x in items]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(fn, items))


@dataclass
class PreparedMolecule:
    coords: np.ndarray        # (n_atoms, 3) float32, centred, sorted by element
    counts: np.ndarray        # (nE,) int64 atoms per element of the pipe's element list
    label: str = ""


@dataclass
class PreparedBatch:
    coords: torch.Tensor      # (total atoms, 3) float32 on the GPU
    mol_off: np.ndarray       # (B+1,) int64
    counts: np.ndarray        # (B, nE) int64
    labels: List[str]

    def __len__(self):
        return len(self.labels)


@dataclass
class RawBatch:
    """Molecules as they come (not centred, atoms in any order) on the GPU,
    for the device-side preparation of ``BatchPipe.run_raw``."""
    coords: torch.Tensor      # (total atoms, 3) float32 on the GPU
    elem: torch.Tensor        # (total atoms,) uint8: element index in the pipe's list
    mol_off: np.ndarray       # (B+1,) int64
    labels: List[str]

    def __len__(self):
        return len(self.labels)


def _centred_extent(z: np.ndarray) -> float:
    """Extent of float32 coordinates z after the centring shift
    ``atoms.centralize`` applies (computed in float32 as it does)."""
    if z.size == 0:
        return 0.0
    lo, hi = z.min(), z.max()
    zc = z + (-lo - (hi - lo) / np.float32(2.0))
    return float(zc.max() - zc.min())


def centred_thickness(mol: atm.AtomList) -> float:
    """z extent of ``atoms.centralize(mol)`` (the slab thickness the
    reference sizes its slices from), without centring the other axes."""
    return _centred_extent(np.asarray(mol.coordinates[:, 0], dtype=np.float32))


def element_indices(mol: atm.AtomList, elements: Sequence[int]) -> np.ndarray:
    """(n_atoms,) uint8 index of each atom's element in ``elements``;
    ValueError for an element outside it."""
    z = np.asarray(mol.elements, dtype=np.int64)
    lut = np.full(max(int(z.max(initial=0)), max(elements)) + 1, 255, dtype=np.uint8)
    lut[list(elements)] = np.arange(len(elements), dtype=np.uint8)
    if z.size and (z.min() < 0 or (lut[z] == 255).any()):
        unknown = sorted(set(np.unique(z).tolist()) - set(elements))
        raise ValueError(f"molecule has elements {unknown} outside the pipe's element "
                         f"list {list(elements)}")
    return lut[z]


class RawStaging:
    """Pinned host buffers a raw batch is packed into, and the event of the
    copies last queued from them (waited for before they are refilled)."""

    def __init__(self):
        self.coords = self.elem = self.done = None

    def acquire(self, n_atoms: int) -> "RawStaging":
        if self.done is not None:
            self.done.synchronize()
        if self.coords is None or self.coords.shape[0] < n_atoms:
            cap = max(int(n_atoms), 1)
            self.coords = torch.empty((cap, 3), dtype=torch.float32, pin_memory=True)
            self.elem = torch.empty((cap,), dtype=torch.uint8, pin_memory=True)
            self.done = torch.cuda.Event()
        return self

    def fill(self, mols: Sequence[atm.AtomList], elements: Sequence[int],
             thickness: bool = False):
        """Write the molecules' coordinates (as given) and element indices
        into the buffers, one slice per molecule on a thread pool; returns
        the molecule offsets and, with ``thickness``, each molecule's
        centred z extent (``centred_thickness``) read off the copy."""
        if len(mols) == 0:
            raise ValueError("empty batch")
        if len(elements) > 254:
            raise ValueError("at most 254 elements per pipe")
        sizes = np.array([len(m) for m in mols], dtype=np.int64)
        mol_off = np.zeros(len(mols) + 1, dtype=np.int64)
        np.cumsum(sizes, out=mol_off[1:])
        self.acquire(int(mol_off[-1]))
        ch, eh = self.coords.numpy(), self.elem.numpy()
        thick = [0.0] * len(mols)

        def one(i):
            a, b = int(mol_off[i]), int(mol_off[i + 1])
            eh[a:b] = element_indices(mols[i], elements)
            c = ch[a:b]
            c[...] = np.asarray(mols[i].coordinates, dtype=np.float32).reshape(-1, 3)
            if thickness:
                thick[i] = _centred_extent(c[:, 0])

        pool_map(one, range(len(mols)))
        return mol_off, thick

    def upload(self, mol_off: np.ndarray, labels) -> RawBatch:
        """Queue the copies of the filled buffers on the current stream
        (without waiting for them) and record their event."""
        n, dev = int(mol_off[-1]), device()
        coords = self.coords[:n].to(dev, non_blocking=True)
        elem = self.elem[:n].to(dev, non_blocking=True)
        self.done.record()
        return RawBatch(coords, elem, mol_off, list(labels))


class MixedBatch:
    """Molecules that need different wave sizes, batched by size.

    Each size group runs as one fused launch sequence (``BatchPipe``); the
    group's slice count is the reference's rule (``atoms.determine_box_size``,
    ceil(extent / dz)) applied to the thickest molecule of the group, so
    every molecule of a group is simulated in the same box.  Images come back
    in input order; ``last_groups`` records (size, slice count, indices)."""

    def __init__(self, microscope: TEM, pixel_size: float, slice_thickness: float,
                 elements: Sequence[int] = (1, 6, 7, 8, 16), symmetric_bandlimit: bool = True):
        self.microscope, self.pixel_size, self.dz = microscope, float(pixel_size), float(
            slice_thickness)
        self.elements = sorted(int(z) for z in elements)
        self.symmetric_bandlimit = symmetric_bandlimit
        self._pipes = {}
        self._stages = {}   # wave size -> RawStaging
        self.last_groups = []

    def pipe(self, n: int, n_slices: int) -> "BatchPipe":
        key = (int(n), int(n_slices))
        if key not in self._pipes:
            self._pipes[key] = BatchPipe(self.microscope, self.pixel_size, self.dz, n, n_slices,
                                         self.elements, self.symmetric_bandlimit)
        return self._pipes[key]

    def slice_count(self, mols: Sequence[atm.AtomList], n: int) -> int:
        return self._slices_for(max(float(m.space[0]) for m in mols), n)

    def _slices_for(self, thickest: float, n: int) -> int:
        ns, _, _ = atm.determine_box_size((thickest, 1.0, 1.0),
                                          voxel_size=(self.dz, self.pixel_size, self.pixel_size),
                                          box_size=(None, n, n))
        return max(int(ns), 1)

    def run(self, items: Sequence[Tuple[atm.AtomList, int]], mode: int = 0,
            device_prep: bool = True) -> List[torch.Tensor]:
        """items: (molecule, wave size); returns one (n, n) image (mode 0) or
        complex64 exit wave (mode 1) per item, on the GPU, in input order.

        device_prep: molecules go to the GPU as they come and are centred and
        grouped by element there (``BatchPipe.run_raw``); the host only finds
        each molecule's thickness (its slice count) and element indices.
        Otherwise they are centred and sorted on the host first
        (``BatchPipe.prepare``).  Both give the same images."""
        out: List[torch.Tensor] = [None] * len(items)
        self.last_groups = []
        mols = [m for m, _ in items]
        groups = [(n, [i for i, (_, sz) in enumerate(items) if int(sz) == n])
                  for n in sorted({int(sz) for _, sz in items})]
        if device_prep:
            # each group's host work -- one pass that packs its atoms into
            # pinned staging and reads off every molecule's thickness -- is
            # done just before its launches, so it overlaps the GPU's work on
            # the group before; the group with the fewest atoms goes first
            groups.sort(key=lambda g: sum(len(mols[i]) for i in g[1]))
            for n, idx in groups:
                st = self._stages.setdefault(n, RawStaging())
                group = [mols[i] for i in idx]
                mol_off, thick = st.fill(group, self.elements, thickness=True)
                ns = self._slices_for(max(thick), n)
                p = self.pipe(n, ns)
                res = p._simulate_raw(st.upload(mol_off, [getattr(m, "label", "")
                                                          for m in group]), mode)
                for j, i in enumerate(idx):
                    out[i] = res[j]
                self.last_groups.append((n, ns, idx))
            return out
        # host preparation of every molecule on a thread pool
        mols = pool_map(atm.centralize, mols)
        thick = pool_map(lambda m: float(m.space[0]), mols)
        for n, idx in groups:
            ns = self._slices_for(max(thick[i] for i in idx), n)
            p = self.pipe(n, ns)
            res = p._simulate(p.upload(p.prepare_many([mols[i] for i in idx])), mode)
            for j, i in enumerate(idx):
                out[i] = res[j]
            self.last_groups.append((n, ns, idx))
        return out


class GraphRun:
    """A captured launch sequence of the fused path (``BatchPipe.capture``)."""

    def __init__(self, graph, out: torch.Tensor, coords: torch.Tensor, keep):
        self.graph, self.out, self.coords = graph, out, coords
        self._keep = keep              # device buffers the graph reads

    def replay(self) -> torch.Tensor:
        self.graph.replay()
        return self.out


class BatchPipe:
    """Batched form of ``Pipe`` for molecules whose atoms are uploaded once.

    ``add_water=True`` is the reference's default image (``pipe.py:16``):
    every voxel empty of all elements gets Poisson water
    (``slice_builder_base.py:134-147``).  Each molecule draws from its own
    seed, the value ``MultiSlicesBuilder.add_water`` would take for it:
    given per call, else drawn from numpy's global RNG, one per molecule in
    batch order."""

    def __init__(self, microscope: TEM, pixel_size: float, slice_thickness: float,
                 roi: int, n_slices: int, elements: Sequence[int] = (1, 6, 7, 8, 16),
                 symmetric_bandlimit: bool = True, add_water: bool = False):
        self.microscope = microscope
        self.add_water = bool(add_water)
        self.pixel_size = float(pixel_size)
        self.dz = float(slice_thickness)
        self.n = int(roi)
        self.n_slices = int(n_slices)
        self.elements = sorted(int(z) for z in elements)
        self.symmetric_bandlimit = symmetric_bandlimit
        if self.add_water and (1 not in self.elements or 8 not in self.elements):
            raise ValueError("add_water needs H (Z=1) and O (Z=8) among the pipe's elements")
        self.builder = MultiSlicesBuilder(self.elements, self.n_slices, self.n, self.n, self.dz,
                                          self.pixel_size, self.pixel_size)
        self.propagator = WavePropagator(self.n, self.n, self.pixel_size, self.pixel_size,
                                         microscope.beam_energy_kev)
        self._bufs = {}

    # -- host preparation (outside any timed region) -----------------------
    def prepare(self, mol: atm.AtomList) -> PreparedMolecule:
        """Centre and sort one molecule by element (host side)."""
        mol = atm.centralize(mol)
        unknown = set(np.unique(mol.elements).tolist()) - set(self.elements)
        if unknown:
            raise ValueError(f"molecule has elements {sorted(unknown)} outside the pipe's "
                             f"element list {self.elements}")
        mol, zs, counts = atm.sort_elements_and_count(mol, self.elements)
        return PreparedMolecule(np.ascontiguousarray(mol.coordinates, dtype=np.float32),
                                np.asarray(counts, dtype=np.int64), mol.label)

    def prepare_many(self, mols: Sequence[atm.AtomList], workers: int = None) \
            -> List[PreparedMolecule]:
        """``prepare`` of every molecule, on a thread pool (``pool_map``)."""
        return pool_map(self.prepare, mols, workers)

    def upload(self, mols: Sequence[PreparedMolecule]) -> PreparedBatch:
        """Concatenate prepared molecules and copy their atoms to the GPU once."""
        if len(mols) == 0:
            raise ValueError("empty batch")
        sizes = np.array([m.coords.shape[0] for m in mols], dtype=np.int64)
        mol_off = np.zeros(len(mols) + 1, dtype=np.int64)
        np.cumsum(sizes, out=mol_off[1:])
        coords = np.concatenate([m.coords for m in mols]) if mol_off[-1] else \
            np.zeros((0, 3), np.float32)
        counts = np.stack([m.counts for m in mols]).astype(np.int64)
        return PreparedBatch(as_device(coords, torch.float32, "coordinates"), mol_off,
                             np.ascontiguousarray(counts), [m.label for m in mols])

    def element_indices(self, mol: atm.AtomList) -> np.ndarray:
        """(n_atoms,) uint8 index of each atom's element in the pipe's list."""
        return element_indices(mol, self.elements)

    def pack_raw(self, mols: Sequence[atm.AtomList], elem=None):
        """A raw batch as plain host arrays -- (coordinates (n, 3) float32 as
        given, element indices (n,) uint8, molecule offsets, labels) -- for
        a producer thread that packs while the GPU simulates (``BatchSim``);
        ``upload_packed`` copies them to the GPU.  ``elem``: the molecules'
        element indices when the caller has them already."""
        if len(mols) == 0:
            raise ValueError("empty batch")
        if len(self.elements) > 254:
            raise ValueError("at most 254 elements per pipe")
        sizes = np.array([len(m) for m in mols], dtype=np.int64)
        mol_off = np.zeros(len(mols) + 1, dtype=np.int64)
        np.cumsum(sizes, out=mol_off[1:])
        if elem is None:
            elem = pool_map(self.element_indices, mols)
        coords = np.concatenate([np.asarray(m.coordinates, dtype=np.float32).reshape(-1, 3)
                                 for m in mols]) if mol_off[-1] else np.zeros((0, 3), np.float32)
        elem = np.concatenate(elem) if mol_off[-1] else np.zeros(0, np.uint8)
        return coords, np.ascontiguousarray(elem), mol_off, [getattr(m, "label", "") for m in mols]

    def upload_packed(self, packed) -> RawBatch:
        """Copy a ``pack_raw`` batch to the GPU through this pipe's pinned
        staging buffers, without waiting for the copy."""
        coords, elem, mol_off, labels = packed
        n = int(mol_off[-1])
        st = self._raw_staging().acquire(n)
        np.copyto(st.coords.numpy()[:n], np.asarray(coords, dtype=np.float32).reshape(-1, 3))
        np.copyto(st.elem.numpy()[:n], np.asarray(elem, dtype=np.uint8))
        return st.upload(mol_off, labels)

    def upload_raw(self, mols: Sequence[atm.AtomList]) -> RawBatch:
        """Molecules as they come, copied to the GPU once: coordinates as
        given (not centred, atoms in any order) and per-atom element indices;
        ``run_raw`` prepares them on the device.

        The host writes both arrays straight into pinned staging buffers
        (one slice per molecule, on a thread pool) and queues their copies
        on the current stream without waiting for them, so the packing of
        the next batch overlaps this one's simulation.  The buffers belong
        to this pipe; the next call waits for these copies before refilling
        them."""
        st = self._raw_staging()
        mol_off, _ = st.fill(mols, self.elements)
        return st.upload(mol_off, [getattr(m, "label", "") for m in mols])

    def _raw_staging(self) -> RawStaging:
        st = getattr(self, "_stage", None)
        if st is None:
            st = self._stage = RawStaging()
        return st

    def upload_orientations(self, mol: atm.AtomList, quats: np.ndarray) -> PreparedBatch:
        """One molecule in len(quats) orientations, rotated and centred on the
        GPU: the batch that ``[prepare(rotated(mol, q)) for q in quats]``
        would upload, with the rotation done in float32 on the device
        (``synth.rotated`` rotates in float64 on the host, so coordinates
        agree to float32 rounding).  The rotation is about the bounding-box
        centre with the reference's matrix convention
        (``utils.rot.get_rotation_mattrices``)."""
        from .utils.rot import get_rotation_mattrices
        quats = np.atleast_2d(np.asarray(quats, dtype=np.float64))
        B = quats.shape[0]
        base = self.prepare(mol)          # sorted by element; rotation keeps the order
        n = base.coords.shape[0]
        if n == 0:
            raise ValueError("molecule has no atoms")
        c = base.coords
        centre = c.min(axis=0) + (c.max(axis=0) - c.min(axis=0)) / 2.0
        rot = torch.as_tensor(get_rotation_mattrices(quats).astype(np.float32).reshape(B, 9))
        dev = device()
        out = torch.empty((B * n, 3), dtype=torch.float32, device=dev)
        ws = torch.empty(max(int(hip_ext.raw("emsim_hip_orient_workspace_size")(B)), 1),
                         dtype=torch.uint8, device=dev)
        base_t = as_device(base.coords, torch.float32, "coordinates")
        rot_t = rot.to(dev)
        hip_ext.call("emsim_hip_orient_batch", ptr(base_t), n, ptr(rot_t), float(centre[0]),
                     float(centre[1]), float(centre[2]), B, ptr(out), ptr(ws), ws.numel(),
                     stream_handle())
        mol_off = np.arange(B + 1, dtype=np.int64) * n
        counts = np.tile(base.counts[None, :], (B, 1)).astype(np.int64)
        return PreparedBatch(out, mol_off, np.ascontiguousarray(counts), [base.label] * B)

    def _buffer(self, name, shape, dtype):
        t = self._bufs.get(name)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = torch.empty(shape, dtype=dtype, device=device())
            self._bufs[name] = t
        return t

    # -- fused GPU path --------------------------------------------------------
    def water_seeds(self, batch: PreparedBatch, seeds=None) -> np.ndarray:
        """(B,) uint64 water seeds: the given ones, else one draw per molecule
        from numpy's global RNG (as ``MultiSlicesBuilder.add_water`` draws)."""
        if seeds is None:
            seeds = [int(np.random.randint(0, 2 ** 63 - 1, dtype=np.int64))
                     for _ in range(len(batch))]
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64).reshape(-1))
        if seeds.size != len(batch):
            raise ValueError(f"need one water seed per molecule ({len(batch)}), got {seeds.size}")
        return seeds

    def _simulate(self, batch: PreparedBatch, mode: int, slices_out=None,
                  seeds=None) -> torch.Tensor:
        from ctypes import c_uint64
        from .physics import water_num_dens
        B = len(batch)
        n_atoms = batch.coords.shape[0]
        if mode == 0:
            out = torch.empty((B, self.n, self.n), dtype=torch.float32, device=device())
        else:
            out = torch.empty((B, self.n, self.n), dtype=torch.complex64, device=device())
        water = "_water" if self.add_water else ""
        ws_bytes = hip_ext.raw(f"emsim_hip_simulate{water}_workspace_size")(
            self.builder.backend.h, B, n_atoms, mode)
        ws = self._buffer("simulate_ws", (max(ws_bytes, 1),), torch.uint8)
        mic = self.microscope
        psi0 = self.propagator.init_wave_value(mic.electron_dose)
        head = (self.builder.backend.h, self.propagator.backend.h, ptr(batch.coords), n_atoms,
                batch.mol_off.ctypes.data_as(POINTER(c_int64)),
                batch.counts.ctypes.data_as(POINTER(c_int64)), B, float(psi0), float(mic.cs_mm),
                float(mic.defocus), float(mic.aperture), 1 if self.symmetric_bandlimit else 0)
        tail = (ptr(out), None if slices_out is None else ptr(slices_out), mode, ptr(ws),
                ws.numel(), stream_handle())
        if not self.add_water:
            hip_ext.call("emsim_hip_simulate", *head, *tail)
            return out
        seeds = self.water_seeds(batch, seeds)
        voxel_water = water_num_dens * self.pixel_size * self.pixel_size * self.dz
        hip_ext.call("emsim_hip_simulate_water", *head, self.elements.index(1),
                     self.elements.index(8), float(voxel_water),
                     seeds.ctypes.data_as(POINTER(c_uint64)), *tail)
        return out

    def _simulate_raw(self, raw: RawBatch, mode: int) -> torch.Tensor:
        if self.add_water:
            raise ValueError("run_raw: the water fill needs host-prepared molecules (run)")
        B, n_atoms = len(raw), raw.coords.shape[0]
        dtype = torch.float32 if mode == 0 else torch.complex64
        out = torch.empty((B, self.n, self.n), dtype=dtype, device=device())
        ws_bytes = hip_ext.raw("emsim_hip_simulate_raw_workspace_size")(
            self.builder.backend.h, B, n_atoms, mode)
        ws = self._buffer("simulate_ws", (max(ws_bytes, 1),), torch.uint8)
        mic = self.microscope
        psi0 = self.propagator.init_wave_value(mic.electron_dose)
        hip_ext.call("emsim_hip_simulate_raw", self.builder.backend.h, self.propagator.backend.h,
                     ptr(raw.coords), ptr(raw.elem), n_atoms,
                     raw.mol_off.ctypes.data_as(POINTER(c_int64)), B, float(psi0),
                     float(mic.cs_mm), float(mic.defocus), float(mic.aperture),
                     1 if self.symmetric_bandlimit else 0, ptr(out), None, mode, ptr(ws),
                     ws.numel(), stream_handle())
        return out

    def run_raw(self, raw: RawBatch) -> torch.Tensor:
        """(B, n, n) float32 images of molecules uploaded as they come
        (``upload_raw``): centred and grouped by element on the device."""
        return self._simulate_raw(raw, 0)

    def run(self, batch: PreparedBatch, seeds=None) -> torch.Tensor:
        """(B, n, n) float32 images on the GPU (``seeds``: water seeds, used
        when the pipe adds water)."""
        return self._simulate(batch, 0, seeds=seeds)

    # -- captured HIP graph -----------------------------------------------------
    def element_offsets(self, batch: PreparedBatch) -> np.ndarray:
        """(B, nE + 1) absolute atom offsets of each molecule's element groups."""
        B, nE = len(batch), len(self.elements)
        counts = np.asarray(batch.counts, dtype=np.int64).reshape(B, nE)
        if (counts < 0).any():
            raise ValueError("negative element count")
        off = np.empty((B, nE + 1), dtype=np.int64)
        off[:, 0] = batch.mol_off[:-1]
        off[:, 1:] = batch.mol_off[:-1, None] + np.cumsum(counts, axis=1)
        if (off[:, -1] > batch.mol_off[1:]).any() or (np.diff(batch.mol_off) < 0).any():
            raise ValueError("element counts exceed a molecule's atoms")
        return off

    def capture(self, batch: PreparedBatch, mode: int = 0) -> "GraphRun":
        """Record the whole fused launch sequence for this batch layout in a
        HIP graph (torch.cuda.CUDAGraph on ROCm).  Replays read the atoms from
        ``GraphRun.coords`` (overwrite it in place with molecules of the same
        per-element counts) and write ``GraphRun.out``.  Per-kernel timing
        must be off while capturing."""
        if self.add_water:
            raise ValueError("capture: the water fill is not available in captured graphs")
        B, n_atoms = len(batch), int(batch.coords.shape[0])
        dev = device()
        mol_off = torch.as_tensor(batch.mol_off, dtype=torch.int64).to(dev)
        elem_off = torch.as_tensor(self.element_offsets(batch)).to(dev)
        coords = batch.coords.clone()
        dtype = torch.float32 if mode == 0 else torch.complex64
        out = torch.empty((B, self.n, self.n), dtype=dtype, device=dev)
        ws_bytes = hip_ext.raw("emsim_hip_simulate_workspace_size")(
            self.builder.backend.h, B, n_atoms, mode)
        ws = torch.empty(max(int(ws_bytes), 1), dtype=torch.uint8, device=dev)

        def launch():
            mic = self.microscope
            psi0 = self.propagator.init_wave_value(mic.electron_dose)
            hip_ext.call("emsim_hip_simulate_dev", self.builder.backend.h,
                         self.propagator.backend.h, ptr(coords), n_atoms, ptr(mol_off),
                         ptr(elem_off), B, float(psi0), float(mic.cs_mm), float(mic.defocus),
                         float(mic.aperture), 1 if self.symmetric_bandlimit else 0, ptr(out),
                         None, mode, ptr(ws), ws.numel(), stream_handle())

        launch()                       # sets up the library's per-device tables
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            launch()
        return GraphRun(graph, out, coords, (mol_off, elem_off, ws))

    def exit_waves(self, batch: PreparedBatch, seeds=None) -> torch.Tensor:
        """(B, n, n) complex64 exit waves on the GPU."""
        return self._simulate(batch, 1, seeds=seeds)

    def exit_waves_and_slices(self, batch: PreparedBatch, slices: torch.Tensor = None,
                              seeds=None):
        """Exit waves plus the projected potential slices the fused path built on the
        way, (B, n_slices, n, n) float32 (written into ``slices`` if given)."""
        shape = (len(batch), self.n_slices, self.n, self.n)
        if slices is None:
            slices = torch.empty(shape, dtype=torch.float32, device=device())
        elif tuple(slices.shape) != shape or slices.dtype != torch.float32 or \
                not slices.is_contiguous():
            raise ValueError(f"slices must be a contiguous float32 tensor of shape {shape}")
        return self._simulate(batch, 1, slices, seeds=seeds), slices

    # -- reference-shaped GPU path (cross-check) --------------------------------
    def build_slices(self, batch: PreparedBatch, seeds=None) -> torch.Tensor:
        """Potential slices of every molecule through the reference-shaped
        builder calls (bin, add_water when the pipe adds water, make)."""
        nE, nS, n = len(self.elements), self.n_slices, self.n
        V = self._buffer("V", (len(batch), nS, n, n), torch.float32)
        hist = self._buffer("hist", (nE, nS, n, n), torch.float32)
        ws = self._buffer("ws", (self.builder.backend.slices_workspace_bytes(),), torch.uint8)
        if self.add_water:
            seeds = self.water_seeds(batch, seeds)
        for b in range(len(batch)):
            lo, hi = batch.mol_off[b], batch.mol_off[b + 1]
            self.builder.backend.bin_atoms(batch.coords[lo:hi], batch.counts[b], out=hist)
            if self.add_water:
                self.builder.add_water(hist, seed=int(seeds[b]))
            self.builder.backend.make_slices(hist, self.symmetric_bandlimit, out=V[b], ws=ws)
        return V

    # -- weak-phase-object images -------------------------------------------------
    def _one_slice_builder(self):
        from .backend.slice_builder_hip import OneSliceBuilder
        if getattr(self, "_one", None) is None:
            self._one = OneSliceBuilder(self.elements, self.n, self.n, self.pixel_size,
                                        self.pixel_size)
        return self._one.backend

    def run_wpo(self, batch: PreparedBatch) -> torch.Tensor:
        """Weak-phase-object images of the whole batch (``pipe.py:57-64,73-77``)
        in one fused launch sequence: each molecule's 2-D projected potential
        (one-slice binning, ``slice_builder_base.py:202-225``, formed from the
        atom buckets like the multislice path's), transmission, the dz = 0
        propagation (the 2/3 band-limit alone), lens and |psi|^2."""
        one = self._one_slice_builder()
        B, n_atoms = len(batch), batch.coords.shape[0]
        out = torch.empty((B, self.n, self.n), dtype=torch.float32, device=device())
        ws_bytes = hip_ext.raw("emsim_hip_simulate_workspace_size")(one.h, B, n_atoms, 0)
        ws = self._buffer("wpo_ws", (max(ws_bytes, 1),), torch.uint8)
        mic = self.microscope
        psi0 = self.propagator.init_wave_value(mic.electron_dose)
        hip_ext.call("emsim_hip_simulate", one.h, self.propagator.backend.h, ptr(batch.coords),
                     n_atoms, batch.mol_off.ctypes.data_as(POINTER(c_int64)),
                     batch.counts.ctypes.data_as(POINTER(c_int64)), B, float(psi0),
                     float(mic.cs_mm), float(mic.defocus), float(mic.aperture),
                     1 if self.symmetric_bandlimit else 0, ptr(out), None, 0, ptr(ws),
                     ws.numel(), stream_handle())
        return out

    def run_wpo_unfused(self, batch: PreparedBatch) -> torch.Tensor:
        """The same images through the reference-shaped steps, molecule by
        molecule: one-slice histogram, make_one_slice, then one batched
        transmission, dz = 0 propagation, lens and |psi|^2."""
        n, nE, B = self.n, len(self.elements), len(batch)
        one = self._one_slice_builder()
        V = self._buffer("V_wpo", (B, n, n), torch.float32)
        hist = self._buffer("hist_wpo", (nE, n, n), torch.float32)
        ws = self._buffer("ws_wpo", (max(one.slices_workspace_bytes(), 1),), torch.uint8)
        for b in range(B):
            lo, hi = batch.mol_off[b], batch.mol_off[b + 1]
            one.bin_atoms(batch.coords[lo:hi], batch.counts[b], out=hist)
            one.make_slices(hist, self.symmetric_bandlimit, out=V[b], ws=ws)
        mic = self.microscope
        w0 = self.propagator.init_wave(mic.electron_dose, batch=B)
        waves = self.propagator.singleslice_propagate(w0, V, 0.0)
        return self.propagator.lens_intensity(waves, mic.cs_mm, mic.defocus, mic.aperture,
                                              overwrite_wave=True)

    def run_unfused(self, batch: PreparedBatch, seeds=None) -> torch.Tensor:
        V = self.build_slices(batch, seeds)
        w0 = self.propagator.init_wave(self.microscope.electron_dose, batch=len(batch))
        waves = self.propagator.multislice_propagate(w0, V, self.dz)
        mic = self.microscope
        return self.propagator.lens_intensity(waves, mic.cs_mm, mic.defocus, mic.aperture,
                                              overwrite_wave=True)


def batched(items, size: int):
    """Consecutive lists of ``size`` items (the last may be shorter)."""
    if size < 1:
        raise ValueError("batch size must be >= 1")
    chunk = []
    for item in items:
        chunk.append(item)
        if len(chunk) == size:
            yield chunk
            chunk = []
    if chunk:
        yield chunk


class BatchSim:
    """Device-resident form of the reference's simulation loop
    (``emsim/simulator.py:21-91``).

    A host thread groups the molecules of ``mols`` into batches of ``batch``
    and packs each batch's atoms with their element indices
    (``BatchPipe.pack_raw``); the calling thread uploads each batch once and
    runs ``BatchPipe.run_raw`` on the current stream, which centres the
    molecules and groups their atoms by element on the device.  With
    ``device_prep=False`` (and always for a pipe that adds water) the host
    thread centres and sorts the molecules (``BatchPipe.prepare``) and the
    batch runs through ``BatchPipe.run``; the images are the same.
    Each batch's images are copied to the host on a second stream (default
    priority) into a ring of ``depth`` pinned buffers, so the copy of batch
    k and the ``result_handler`` calls for the batch before it overlap the
    simulation of batch k + 1.  ``result_handler(image, label)`` receives a
    float32 numpy array of its own per molecule, in input order.  As in the
    reference, a molecule (or batch) that fails is logged and skipped.
    """

    def __init__(self, pipe: "BatchPipe", mols, result_handler, batch: int = 64,
                 depth: int = 2, device_prep: bool = True):
        if depth < 1:
            raise ValueError("depth must be >= 1")
        self.pipe, self.mols, self.result_handler = pipe, mols, result_handler
        self.device_prep = bool(device_prep) and not pipe.add_water
        self.batch, self.depth = int(batch), int(depth)
        self.delivered = 0
        self.failed = 0

    def _producer(self, q, stop):
        import logging
        log = logging.getLogger(__name__)
        try:
            def attempt(mol):
                try:
                    if self.device_prep:   # refuses unknown elements
                        return mol, self.pipe.element_indices(mol)
                    return self.pipe.prepare(mol)
                except Exception as e:   # reference: log the failed image, go on
                    return e

            for chunk in batched(self.mols, self.batch):
                prepared = []
                for mol, res in zip(chunk, pool_map(attempt, chunk)):
                    if isinstance(res, Exception):
                        log.warning(f"molecule {getattr(mol, 'label', '')!r} skipped: {res}")
                        self.failed += 1
                    else:
                        prepared.append(res)
                if prepared and self.device_prep:
                    q.put(self.pipe.pack_raw([m for m, _ in prepared],
                                             [e for _, e in prepared]))
                elif prepared:
                    q.put(prepared)
                if stop.is_set():
                    break
        except Exception as e:   # an iterator that raises ends the stream
            log.warning(f"molecule iterator failed with {type(e)}: {e}")
        q.put(None)

    def run(self) -> int:
        """Simulate every molecule; returns the number of images delivered."""
        import logging
        from collections import deque
        from queue import Queue
        from threading import Event, Thread
        log = logging.getLogger(__name__)
        n = self.pipe.n
        q, stop = Queue(maxsize=2), Event()
        producer = Thread(target=self._producer, args=(q, stop), daemon=True)
        producer.start()
        copy_stream = torch.cuda.Stream()
        ring = [torch.empty((self.batch, n, n), dtype=torch.float32, pin_memory=True)
                for _ in range(self.depth)]
        pending = deque()   # (slot, copy event, device images kept alive, labels)
        slot = 0

        def deliver():
            s, ev, _, labels = pending.popleft()
            ev.synchronize()
            host = ring[s].numpy()
            for i, label in enumerate(labels):
                self.result_handler(host[i].copy(), label)
                self.delivered += 1

        try:
            while True:
                prepared = q.get()
                if prepared is None:
                    break
                labels = prepared[3] if self.device_prep else [m.label for m in prepared]
                try:
                    if self.device_prep:
                        images = self.pipe.run_raw(self.pipe.upload_packed(prepared))
                    else:
                        images = self.pipe.run(self.pipe.upload(prepared))
                except Exception as e:
                    log.warning(f"batch of {len(labels)} molecules failed: {e}")
                    self.failed += len(labels)
                    continue
                done = torch.cuda.Event()
                done.record()
                if len(pending) == self.depth:
                    deliver()   # frees the ring slot about to be reused
                copy_stream.wait_event(done)
                with torch.cuda.stream(copy_stream):
                    ring[slot][:len(labels)].copy_(images, non_blocking=True)
                    copied = torch.cuda.Event()
                    copied.record(copy_stream)
                images.record_stream(copy_stream)
                pending.append((slot, copied, images, labels))
                slot = (slot + 1) % self.depth
            while pending:
                deliver()
        finally:
            stop.set()
            while producer.is_alive():   # unblock a producer waiting on a full queue
                try:
                    q.get_nowait()
                except Exception:
                    pass
                producer.join(timeout=0.05)
        return self.delivered
