"""The record every material generator yields (materials/materialmodel.py:11-29)."""
from dataclasses import dataclass

import numpy as np


@dataclass
class MaterialModel:
    """Relative permittivity and conductivity [S/m] of the material volume.

    Both arrays have shape (size_z_material, size_y, size_x), C order, x the
    fastest axis.  ``name`` names the results file; ``metadata`` is stored
    with the results.
    """
    relative_permittivity: np.ndarray
    conductivity: np.ndarray
    cell_size: float
    name: str = 'results'
    metadata: dict = None
