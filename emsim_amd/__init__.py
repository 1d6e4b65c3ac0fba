"""emsim_amd: an MI355X (gfx950) HIP multislice backend behind emsim's API.

Modules mirror the reference package's hot path: ``config`` (backend
registry, ``set_backend('hip')``), ``pot`` (slice builders), ``wave``
(propagator frontend), ``tem``, ``pipe``, ``simulator``, ``atoms``,
``elem``, ``physics`` and ``utils``.  The GPU code lives in ``csrc/`` and is
loaded through ``backend/hip_ext``.
"""
import logging as _logging

from . import config
from . import utils
from . import physics
from . import elem
from . import atoms
from . import tem
from . import pipe
from . import simulator
from . import pot
from . import wave

global_logger = _logging.getLogger(__name__)
if not global_logger.handlers:
    _ch = _logging.StreamHandler()
    _ch.setLevel(_logging.INFO)
    _ch.setFormatter(_logging.Formatter('%(asctime)s::%(levelname)s::%(message)s'))
    global_logger.addHandler(_ch)
    global_logger.setLevel(_logging.INFO)
