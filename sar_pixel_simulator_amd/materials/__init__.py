from .materialmodel import MaterialModel
from .materials import (compute_material_arrays, all_vacuum, flat_surface,
                        rough_dielectric, fill_below_surface)
from .roughsurface import gaussian_correlated_surface, two_layer_gaussian_correlated_surface
from . import dielectricmodels, device, roughsurface, soilmodels, synthetic

__all__ = ['MaterialModel', 'compute_material_arrays', 'all_vacuum',
           'flat_surface', 'rough_dielectric', 'fill_below_surface',
           'gaussian_correlated_surface', 'two_layer_gaussian_correlated_surface',
           'dielectricmodels', 'device', 'roughsurface', 'soilmodels', 'synthetic']
