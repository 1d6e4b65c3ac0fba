from . import rot
from . import array
