"""Per-image pipeline (``emsim/pipe.py:9-91``).

``Pipe`` keeps the reference's constructor, properties and methods:
centralize -> build_multi_slices -> init_wave -> multislice -> lens -> |psi|^2.
Two reference behaviours are kept on purpose: ``n_slices`` is stored but
not passed to the slice builder (``pipe.py:48-52``), so the slice count
follows from the molecule's thickness, and ``add_water`` defaults to True.
Propagators are looked up per call as in the reference (``pipe.py:31-34``);
the hip backend caches their device state, so this costs no re-planning.
The intensity is formed by the fused lens + |psi|^2 kernel.
"""
from typing import Optional, Tuple, Union

from . import atoms as atm
from . import pot, tem, wave


class Pipe(object):
    def __init__(self, microscope: tem.TEM, resolution: float, slice_thickness: float,
                 roi: Optional[Union[int, Tuple[int, int]]] = None,
                 n_slices: Optional[int] = None, add_water: bool = True,
                 upto_exit_wave: bool = False):
        self._resolution = resolution
        self._pixel_size = 0.5 * resolution
        self.microscope = microscope
        self.slice_thickness = slice_thickness
        self.add_water = add_water
        self.upto_exit_wave = upto_exit_wave
        self.roi = (roi, roi) if type(roi) is int else roi
        self.n_slices = n_slices

    @property
    def wave_propagator(self):
        return wave.get_raw_wave_propagator(self.roi, self._pixel_size,
                                            self.microscope.beam_energy_kev)

    @property
    def resolution(self):
        return self._resolution

    @resolution.setter
    def resolution(self, res):
        self._pixel_size = 0.5 * res
        self._resolution = res

    def get_exit_wave(self, mol):
        mol = atm.centralize(mol)
        slices = pot.build_multi_slices(mol, dz=self.slice_thickness,
                                        pixel_size=self._pixel_size, lateral_size=self.roi,
                                        add_water=self.add_water)
        prop = self.wave_propagator
        init_wave = prop.init_wave(self.microscope.electron_dose)
        return prop.multislice_propagate(init_wave, slices, self.slice_thickness)

    def get_exit_wave_wpo(self, mol):
        mol = atm.centralize(mol)
        aslice = pot.build_one_slice(mol, pixel_size=self._pixel_size, lateral_size=self.roi)
        prop = self.wave_propagator
        init_wave = prop.init_wave(self.microscope.electron_dose)
        return prop.singleslice_propagate(init_wave, aslice, dz=0)

    def lens_propagate(self, exit_wave):
        return self.wave_propagator.lens_propagate(exit_wave, self.microscope.cs_mm,
                                                   self.microscope.defocus,
                                                   self.microscope.aperture)

    def _image(self, exit_wave):
        return self.wave_propagator.lens_intensity(exit_wave, self.microscope.cs_mm,
                                                   self.microscope.defocus,
                                                   self.microscope.aperture,
                                                   overwrite_wave=True)

    def run_wpo(self, mol):
        return self._image(self.get_exit_wave_wpo(mol))

    def run(self, mol):
        return self._image(self.get_exit_wave(mol))

    def __repr__(self):
        return (f"{{ImagePipe | resolution = {self._resolution}Angstrom, "
                f"slice_thickness = {self.slice_thickness}Angstrom, "
                f"roi = {self.roi}, "
                f"add_water = {self.add_water}, "
                f"microscope = {self.microscope}}}")
