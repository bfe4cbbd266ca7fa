"""MI355X-native zodiacal emission: a drop-in for ``zodipy.Model.evaluate``.

Exports the same names as the reference package (zodipy/__init__.py:1-9):
``Model``, ``model_registry`` and ``grid_number_density``.
"""

from zodipy_amd import units
from zodipy_amd.coords import SkyCoord, Time
from zodipy_amd.model import Model
from zodipy_amd.model_registry import model_registry
from zodipy_amd.number_density import grid_number_density

__all__ = (
    "Model",
    "SkyCoord",
    "Time",
    "grid_number_density",
    "model_registry",
    "units",
)
