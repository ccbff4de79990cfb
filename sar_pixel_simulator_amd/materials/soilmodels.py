"""Soil material generators (materials/soilmodels.py:16-298).

Each generator yields MaterialModel's of (size_z_material, size_y, size_x)
permittivity and conductivity: vacuum above one or two Gaussian-correlated
rough surfaces, soil below them with Hallikainen permittivity.  Surface
heights are rint(h / dx + mean depth in cells) as int32 and a cell is soil
where z <= height (soilmodels.py:43-62).  The fills are vectorised; the
values are those of the reference's loops.

``*_layers`` companions return what the device builder
(``materials.device``) needs for the same model -- the int32 surface maps
and per-layer values -- so that the 3-D arrays are formed in HBM instead of
being shipped from the host.
"""
from typing import Sequence

import numpy as np

from ..parameters import ParameterCollection
from .dielectricmodels import HallikainenSoilModel, soil_eps_sigma
from .materialmodel import MaterialModel
from .roughsurface import gaussian_correlated_surface, two_layer_gaussian_correlated_surface


def _grid(params):
    g = params.geom_params
    return (g.size_x, g.size_y, g.size_z_material, params.processing_params.datatype,
            g.cell_size, params.inc_beam_params.frequency)


def one_layer_surfaces(params, height_rms, corr_len, mean_depth, count, seed, grooved=False,
                       rabus_compat=False):
    """The int32 surface maps (ny, nx) of one_layer_gaussian's ``count``
    realizations (soilmodels.py:52-62)."""
    nx, ny, nz, _, dx, _ = _grid(params)
    if mean_depth > nz * dx:
        raise ValueError('Requested mean soil depth is greater than the max '
                         'height of the simulator space.')
    for i in range(count):
        s = gaussian_correlated_surface(nx, ny, dx, corr_len, corr_len, height_rms,
                                        seed=seed + i, rabus_compat=rabus_compat)
        if grooved:
            s = np.repeat(s[0:1, :], ny, axis=0)
        s /= dx
        s += mean_depth / dx
        yield np.rint(s).astype(np.int32)


def two_layer_surfaces(params, height_rms, corr_len, inter_layer_corr, mean_total_depth,
                       top_layer_height, count, seed, grooved=False, rabus_compat=False):
    """The int32 surface maps (2, ny, nx) of the two-layer generators:
    [0] the lower layer's top, [1] the upper layer's top
    (soilmodels.py:153-166)."""
    nx, ny, nz, _, dx, _ = _grid(params)
    total = mean_total_depth / dx
    top = top_layer_height / dx
    if mean_total_depth > nz * dx:
        raise ValueError('Requested mean soil depth is greater than the max '
                         'height of the simulator space.')
    if top > total:
        raise ValueError('Top layer of soil is thicker than entire soil model')
    for i in range(count):
        s = two_layer_gaussian_correlated_surface(nx, ny, dx, corr_len, corr_len, height_rms,
                                                  inter_layer_corr, seed + i,
                                                  rabus_compat=rabus_compat)
        if grooved:
            s = np.repeat(s[:, 0:1, :], ny, axis=1)
        s /= dx
        s += total
        s[0] -= top
        yield np.rint(s).astype(np.int32)


def _below(nz, surf):
    return np.arange(nz).reshape(nz, 1, 1) <= surf[np.newaxis]


def one_layer_gaussian(params: ParameterCollection, moistures: Sequence[float], sand_frac: float,
                       clay_frac: float, height_rms: float, corr_len: float, mean_depth: float,
                       count: int, seed: int, grooved: bool = False, rabus_compat: bool = False):
    """Soil of constant Hallikainen permittivity under one rough surface, for
    every realization and moisture (soilmodels.py:16-90)."""
    nx, ny, nz, dtype, dx, freq = _grid(params)
    model = HallikainenSoilModel()
    for i, surf in enumerate(one_layer_surfaces(params, height_rms, corr_len, mean_depth, count,
                                                seed, grooved, rabus_compat)):
        below = _below(nz, surf)
        for mv in moistures:
            eps_r, cond = soil_eps_sigma(model.get_epsilon(mv, sand_frac, clay_frac, freq), freq)
            meta = {'moisture': mv, 'sand_frac': sand_frac, 'clay_frac': clay_frac,
                    'epsilon_r': eps_r, 'conductivity': cond, 'height_rms': height_rms,
                    'mean_depth': mean_depth, 'mean_depth_cells': mean_depth / dx,
                    'corr_len': corr_len}
            eps = np.ones((nz, ny, nx), dtype=dtype)
            sig = np.zeros((nz, ny, nx), dtype=dtype)
            eps[below] = eps_r
            sig[below] = cond
            yield MaterialModel(eps, sig, dx, f'mv_{mv:.2f}_num_{i}', meta)


def _two_layer_meta(top_moisture, mv, sand_frac, clay_frac, height_rms, mean_total_depth,
                    top_layer_height, inter_layer_corr, corr_len, dx):
    return {'top_moisture': top_moisture, 'bottom_moisture': mv, 'sand_frac': sand_frac,
            'clay_frac': clay_frac, 'height_rms': height_rms,
            'mean_total_depth': mean_total_depth, 'total_depth_cells': mean_total_depth / dx,
            'top_layer_height': top_layer_height, 'top_layer_cells': top_layer_height / dx,
            'inter_layer_corr': inter_layer_corr, 'corr_len': corr_len}


def two_layer_gaussian(params: ParameterCollection, top_moisture: float,
                       bottom_moistures: Sequence[float], sand_frac: float, clay_frac: float,
                       height_rms: float, corr_len: float, inter_layer_corr: float,
                       mean_total_depth: float, top_layer_height: float, count: int, seed: int,
                       grooved: bool = False, rabus_compat: bool = False):
    """Two soil layers of constant Hallikainen permittivity: the bottom
    moisture where z <= surface 0, the top moisture where surface 0 < z <=
    surface 1 (soilmodels.py:93-197)."""
    nx, ny, nz, dtype, dx, freq = _grid(params)
    model = HallikainenSoilModel()
    e_top, s_top = soil_eps_sigma(model.get_epsilon(top_moisture, sand_frac, clay_frac, freq), freq)
    for i, surf in enumerate(two_layer_surfaces(params, height_rms, corr_len, inter_layer_corr,
                                                mean_total_depth, top_layer_height, count, seed,
                                                grooved, rabus_compat)):
        lo = _below(nz, surf[0])
        hi = _below(nz, surf[1]) & ~lo
        for mv in bottom_moistures:
            e_bot, s_bot = soil_eps_sigma(model.get_epsilon(mv, sand_frac, clay_frac, freq), freq)
            meta = _two_layer_meta(top_moisture, mv, sand_frac, clay_frac, height_rms,
                                   mean_total_depth, top_layer_height, inter_layer_corr,
                                   corr_len, dx)
            meta.update({'epsilon_r': np.asarray([e_bot, e_top]),
                         'conductivity': np.asarray([s_bot, s_top])})
            eps = np.ones((nz, ny, nx), dtype=dtype)
            sig = np.zeros((nz, ny, nx), dtype=dtype)
            eps[lo] = e_bot
            sig[lo] = s_bot
            eps[hi] = e_top
            sig[hi] = s_top
            yield MaterialModel(eps, sig, dx, f'mv_{mv:.2f}_num_{i}', meta)


def gradient_moisture(surf, nz, mv_bottom, mv_top):
    """Per-cell moisture of the gradient layer, ((mv_b (h1 - z)) + (mv_t
    (z - h0))) / (h1 - h0), on the cells surface 0 < z <= surface 1
    (soilmodels.py:233-247); NaN elsewhere."""
    h0 = surf[0][np.newaxis].astype(np.int64)
    h1 = surf[1][np.newaxis].astype(np.int64)
    z = np.arange(nz, dtype=np.int64).reshape(nz, 1, 1)
    mask = (z > h0) & (z <= h1)
    den = np.where(mask, h1 - h0, 1)
    mv = ((mv_bottom * (h1 - z)) + (mv_top * (z - h0))) / den
    return np.where(mask, mv, np.nan)


def gradient_gaussian(params: ParameterCollection, top_moisture: float,
                      bottom_moistures: Sequence[float], sand_frac: float, clay_frac: float,
                      height_rms: float, corr_len: float, inter_layer_corr: float,
                      mean_total_depth: float, top_layer_height: float, count: int, seed: int,
                      grooved: bool = False, rabus_compat: bool = False):
    """Two soil layers: constant bottom moisture where z <= surface 0, a
    moisture interpolated linearly in z from the bottom to the top moisture
    where surface 0 < z <= surface 1, each cell's permittivity from the
    Hallikainen model (soilmodels.py:200-298)."""
    from .dielectricmodels import hallikainen_soil_model
    nx, ny, nz, dtype, dx, freq = _grid(params)
    for i, surf in enumerate(two_layer_surfaces(params, height_rms, corr_len, inter_layer_corr,
                                                mean_total_depth, top_layer_height, count, seed,
                                                grooved, rabus_compat)):
        lo = _below(nz, surf[0])
        for mv in bottom_moistures:
            e_bot, s_bot = soil_eps_sigma(hallikainen_soil_model(mv, sand_frac, clay_frac, freq),
                                          freq)
            eps = np.ones((nz, ny, nx), dtype=dtype)
            sig = np.zeros((nz, ny, nx), dtype=dtype)
            eps[lo] = e_bot
            sig[lo] = s_bot
            mvz = gradient_moisture(surf, nz, mv, top_moisture)
            cells = np.nonzero(~np.isnan(mvz))
            if cells[0].size:
                # the model is a polynomial in mv per table frequency, then a
                # frequency interpolation: evaluate it on the distinct moistures
                uniq, inv = np.unique(mvz[cells], return_inverse=True)
                vals = [soil_eps_sigma(hallikainen_soil_model(m, sand_frac, clay_frac, freq), freq)
                        for m in uniq]
                eps[cells] = np.array([v[0] for v in vals])[inv]
                sig[cells] = np.array([v[1] for v in vals])[inv]
            meta = _two_layer_meta(top_moisture, mv, sand_frac, clay_frac, height_rms,
                                   mean_total_depth, top_layer_height, inter_layer_corr,
                                   corr_len, dx)
            yield MaterialModel(eps, sig, dx, f'mv_{mv:.2f}_num_{i}', meta)


__all__ = ['one_layer_gaussian', 'two_layer_gaussian', 'gradient_gaussian',
           'one_layer_surfaces', 'two_layer_surfaces', 'gradient_moisture']
