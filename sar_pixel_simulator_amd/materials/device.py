"""Material models described for the device builder.

A ``MaterialSpec`` carries what ``fdtd_sim_build_materials`` needs to form a
model's Ca, Cb in HBM: up to four layers below int32 surface maps
(constant, random per-cell soil, or a Hallikainen moisture gradient),
Bernoulli voxel inclusions and dielectric cylinders.  The host ships the
(ny, nx) surface maps and the cylinder list, O(nx ny) bytes, instead of two
(nz, ny, nx) float32 volumes.

Constructors:
  * ``one_layer_specs``, ``two_layer_specs``, ``gradient_specs``: the
    reference's soil generators (materials/soilmodels.py:16-298); the device
    builds the same eps_r, sigma as ``soilmodels`` does on the host;
  * ``rough_soil_spec``, ``snow_over_soil_spec``, ``canopy_over_soil_spec``:
    the benchmark workloads' statistics (``synthetic``) with per-cell draws
    made on the device.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np
from scipy.constants import epsilon_0

from .. import _native as native
from ..constants import C0, EPS0_2018
from .dielectricmodels import (FREQ_TABLE, HallikainenSoilModel, hallikainen_soil_model,
                               polynomial_coefficients, soil_eps_sigma, table_bracket)
from .soilmodels import one_layer_surfaces, two_layer_surfaces
from .synthetic import SNOW, SOIL, surface_cells


@dataclass
class Layer:
    """One layer: cells with z <= its surface that no lower layer took."""
    kind: int
    eps: float = 1.0
    sigma: float = 0.0
    eps_std: float = 0.0
    sigma_hi: float = 0.0
    seed: int = 0
    mv_bottom: float = 0.0
    mv_top: float = 0.0
    sand_frac: float = 0.0
    clay_frac: float = 0.0
    freq: float = 0.0

    @staticmethod
    def const(eps, sigma):
        return Layer(native.LAYER_CONST, eps=float(eps), sigma=float(sigma))

    @staticmethod
    def random_soil(seed, eps_mean=SOIL['eps_mean'], eps_std=SOIL['eps_std'],
                    sig_lo=SOIL['sig_lo'], sig_hi=SOIL['sig_hi']):
        return Layer(native.LAYER_RANDOM, eps=eps_mean, sigma=sig_lo, eps_std=eps_std,
                     sigma_hi=sig_hi, seed=int(seed))

    @staticmethod
    def gradient(mv_bottom, mv_top, sand_frac, clay_frac, freq):
        return Layer(native.LAYER_GRADIENT, mv_bottom=float(mv_bottom), mv_top=float(mv_top),
                     sand_frac=float(sand_frac), clay_frac=float(clay_frac), freq=float(freq))

    def to_c(self):
        c = native.FdtdMatLayer()
        c.kind = self.kind
        c.seed = self.seed & 0xffffffff
        c.eps, c.sigma, c.eps_std, c.sigma_hi = self.eps, self.sigma, self.eps_std, self.sigma_hi
        c.mv_bottom, c.mv_top, c.freq = self.mv_bottom, self.mv_top, self.freq
        if self.kind == native.LAYER_GRADIENT:
            c.eps0 = epsilon_0
            j = table_bracket(self.freq)
            c.f0, c.f1 = FREQ_TABLE[j], FREQ_TABLE[j + 1]
            a_re, b_re, c_re, a_im, b_im, c_im = polynomial_coefficients(self.sand_frac,
                                                                         self.clay_frac)
            for r, row in enumerate((j, j + 1)):
                for q, v in enumerate((a_re[row], b_re[row], c_re[row])):
                    c.poly[r][0][q] = v
                for q, v in enumerate((a_im[row], b_im[row], c_im[row])):
                    c.poly[r][1][q] = v
        return c


@dataclass
class Cylinder:
    """Dielectric cylinder: centre and unit axis (z, y, x) [m]."""
    centre: tuple
    axis: tuple
    radius: float
    length: float
    eps: float
    sigma: float


@dataclass
class MaterialSpec:
    """A material model for ``fdtd_sim_build_materials``."""
    nz_mat: int
    layers: list = field(default_factory=list)
    surfaces: np.ndarray = None            # int32 (len(layers), ny, nx)
    cylinders: list = field(default_factory=list)
    vox_z1: int = 0
    vox_frac: float = 0.0
    vox_eps: float = 1.0
    vox_sigma: float = 0.0
    vox_seed: int = 0
    name: str = 'model'
    metadata: dict = field(default_factory=dict)

    def payload_bytes(self):
        """Bytes the host ships for this model."""
        n = ctypes.sizeof(native.FdtdMatSpec)
        if self.surfaces is not None:
            n += self.surfaces.nbytes
        return n + len(self.cylinders) * ctypes.sizeof(native.FdtdMatCylinder)


def cylinder_box(c, dx, shape):
    """Cell box [lo, hi) of a cylinder, as the host generator bounds it
    (``synthetic.canopy_over_soil``)."""
    u = np.asarray(c.axis, dtype=np.float64)
    centre = np.asarray(c.centre, dtype=np.float64)
    half = 0.5 * c.length * np.abs(u) + c.radius * np.sqrt(np.maximum(0.0, 1.0 - u * u))
    lo = np.maximum(np.floor((centre - half) / dx).astype(int), 0)
    hi = np.minimum(np.ceil((centre + half) / dx).astype(int) + 1, list(shape))
    hi = np.maximum(hi, lo)
    return lo, hi


def to_c(spec: MaterialSpec, params):
    """(FdtdMatSpec, surfaces, cylinder array) for ``Sim.build_materials``."""
    g = params.geom_params
    if len(spec.layers) > native.MAX_LAYERS:
        raise ValueError(f'at most {native.MAX_LAYERS} layers')
    c = native.FdtdMatSpec()
    c.nz_mat = spec.nz_mat
    c.nlayers = len(spec.layers)
    c.ncyl = len(spec.cylinders)
    c.vox_z1 = spec.vox_z1
    c.vox_frac, c.vox_eps, c.vox_sigma = spec.vox_frac, spec.vox_eps, spec.vox_sigma
    c.vox_seed = spec.vox_seed & 0xffffffff
    # compute_material_arrays' float32 scalars (numpy casts the Python floats
    # to the array's float32)
    c.dt, c.dx, c.eps0 = np.float32(g.timestep), np.float32(g.cell_size), np.float32(EPS0_2018)
    for k, layer in enumerate(spec.layers):
        c.layer[k] = layer.to_c()
    cyl = None
    if spec.cylinders:
        cyl = (native.FdtdMatCylinder * len(spec.cylinders))()
        shape = (spec.nz_mat, g.size_y, g.size_x)
        for i, cy in enumerate(spec.cylinders):
            lo, hi = cylinder_box(cy, g.cell_size, shape)
            d = cyl[i]
            for q in range(3):
                d.c[q], d.u[q] = cy.centre[q], cy.axis[q]
                d.lo[q], d.hi[q] = int(lo[q]), int(hi[q])
            d.radius, d.length, d.eps, d.sigma = cy.radius, cy.length, cy.eps, cy.sigma
    return c, spec.surfaces, cyl


# ---------------------------------------------------------------------------
# the reference's soil generators as specs
# ---------------------------------------------------------------------------
def one_layer_specs(params, moistures, sand_frac, clay_frac, height_rms, corr_len, mean_depth,
                    count, seed, grooved=False, rabus_compat=False):
    """Specs of ``soilmodels.one_layer_gaussian``'s models, in its order."""
    g = params.geom_params
    freq = params.inc_beam_params.frequency
    model = HallikainenSoilModel()
    for i, surf in enumerate(one_layer_surfaces(params, height_rms, corr_len, mean_depth, count,
                                                seed, grooved, rabus_compat)):
        for mv in moistures:
            eps_r, cond = soil_eps_sigma(model.get_epsilon(mv, sand_frac, clay_frac, freq), freq)
            yield MaterialSpec(g.size_z_material, [Layer.const(eps_r, cond)], surf[np.newaxis],
                               name=f'mv_{mv:.2f}_num_{i}',
                               metadata={'moisture': mv, 'sand_frac': sand_frac,
                                         'clay_frac': clay_frac, 'epsilon_r': eps_r,
                                         'conductivity': cond})


def two_layer_specs(params, top_moisture, bottom_moistures, sand_frac, clay_frac, height_rms,
                    corr_len, inter_layer_corr, mean_total_depth, top_layer_height, count, seed,
                    grooved=False, rabus_compat=False):
    """Specs of ``soilmodels.two_layer_gaussian``'s models, in its order."""
    g = params.geom_params
    freq = params.inc_beam_params.frequency
    model = HallikainenSoilModel()
    top = Layer.const(*soil_eps_sigma(model.get_epsilon(top_moisture, sand_frac, clay_frac, freq),
                                      freq))
    for i, surf in enumerate(two_layer_surfaces(params, height_rms, corr_len, inter_layer_corr,
                                                mean_total_depth, top_layer_height, count, seed,
                                                grooved, rabus_compat)):
        for mv in bottom_moistures:
            bot = Layer.const(*soil_eps_sigma(model.get_epsilon(mv, sand_frac, clay_frac, freq),
                                              freq))
            yield MaterialSpec(g.size_z_material, [bot, top], surf, name=f'mv_{mv:.2f}_num_{i}',
                               metadata={'top_moisture': top_moisture, 'bottom_moisture': mv})


def gradient_specs(params, top_moisture, bottom_moistures, sand_frac, clay_frac, height_rms,
                   corr_len, inter_layer_corr, mean_total_depth, top_layer_height, count, seed,
                   grooved=False, rabus_compat=False):
    """Specs of ``soilmodels.gradient_gaussian``'s models, in its order."""
    g = params.geom_params
    freq = params.inc_beam_params.frequency
    table_bracket(freq)                        # frequency inside the table
    for i, surf in enumerate(two_layer_surfaces(params, height_rms, corr_len, inter_layer_corr,
                                                mean_total_depth, top_layer_height, count, seed,
                                                grooved, rabus_compat)):
        for mv in bottom_moistures:
            bot = Layer.const(*soil_eps_sigma(hallikainen_soil_model(mv, sand_frac, clay_frac,
                                                                     freq), freq))
            grad = Layer.gradient(mv, top_moisture, sand_frac, clay_frac, freq)
            yield MaterialSpec(g.size_z_material, [bot, grad], surf, name=f'mv_{mv:.2f}_num_{i}',
                               metadata={'top_moisture': top_moisture, 'bottom_moisture': mv})


# ---------------------------------------------------------------------------
# the benchmark workloads' statistics, drawn on the device
# ---------------------------------------------------------------------------
def rough_soil_spec(params, mean_cells, seed, name=None):
    """Random soil (eps_r ~ N(8, 0.5^2) clipped at 1, sigma ~ U(0.005, 0.02)
    S/m) under the rough surface of ``synthetic.rough_soil``."""
    g = params.geom_params
    lam = C0 / params.inc_beam_params.frequency
    h = surface_cells(g.size_x, g.size_y, g.cell_size, lam, mean_cells, seed)
    return MaterialSpec(g.size_z_material, [Layer.random_soil(seed + 1)], h[np.newaxis],
                        name=name or f'soil_{seed}', metadata={'seed': seed,
                                                               'mean_cells': mean_cells})


def snow_over_soil_spec(params, soil_cells, snow_cells, seed, name=None):
    """Bernoulli(0.3) ice voxels (eps_r 3.15, sigma 1e-4 S/m) in the planes
    below soil_cells + snow_cells, over random rough soil."""
    spec = rough_soil_spec(params, soil_cells, seed, name or f'snow_{seed}')
    spec.vox_z1 = min(spec.nz_mat, soil_cells + snow_cells)
    spec.vox_frac, spec.vox_eps, spec.vox_sigma = (SNOW['ice_frac'], SNOW['ice_eps'],
                                                   SNOW['ice_sig'])
    spec.vox_seed = seed + 2
    spec.metadata.update({'soil_cells': soil_cells, 'snow_cells': snow_cells})
    return spec


def canopy_cylinders(params, soil_cells, seed, density):
    """The random cylinders of ``synthetic.canopy_over_soil`` (same draws,
    same order)."""
    g = params.geom_params
    lam = C0 / params.inc_beam_params.frequency
    dx, nx, ny, nz = g.cell_size, g.size_x, g.size_y, g.size_z_material
    rng = np.random.default_rng(seed + 3)
    count = int(round(density * (nx * dx / lam) * (ny * dx / lam)))
    out = []
    for _ in range(count):
        r = rng.uniform(0.05, 0.3) * lam
        ln = rng.uniform(0.5, 2.0) * lam
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c = (rng.uniform(soil_cells * dx, nz * dx), rng.uniform(0, ny * dx), rng.uniform(0, nx * dx))
        e_c = rng.uniform(10.0, 25.0)
        s_c = rng.uniform(0.1, 0.5)
        out.append(Cylinder(tuple(c), tuple(u), r, ln, e_c, s_c))
    return out


def canopy_over_soil_spec(params, soil_cells, seed, density, name=None):
    """Random dielectric cylinders over random rough soil."""
    spec = rough_soil_spec(params, soil_cells, seed, name or f'canopy_{seed}')
    spec.cylinders = canopy_cylinders(params, soil_cells, seed, density)
    spec.metadata.update({'soil_cells': soil_cells, 'cylinders': len(spec.cylinders)})
    return spec


__all__ = ['Layer', 'Cylinder', 'MaterialSpec', 'to_c', 'cylinder_box', 'one_layer_specs',
           'two_layer_specs', 'gradient_specs', 'rough_soil_spec', 'snow_over_soil_spec',
           'canopy_cylinders', 'canopy_over_soil_spec']
