"""Material arrays and the reference's stock material generators.

``compute_material_arrays`` turns (eps_r, sigma) into the Yee E-update
coefficients Ca and Cb (Taflove eqn 3.31a,b) exactly as
fdtdsarsim/materials/materials.py:23-68 does: vacuum is appended up to
size_z, permittivity is scaled by the CODATA-2018 eps0 literal, and every
operation runs in the configured float type.
"""
import numpy as np

from ..constants import EPS0_2018
from ..parameters import ParameterCollection
from .materialmodel import MaterialModel
from .roughsurface import gaussian_correlated_surface


def compute_material_arrays(params: ParameterCollection,
                            materials: MaterialModel):
    """Return (Ca, Cb), each (size_z, size_y, size_x) in ``datatype``.

    Ca = (1 - s dt / 2e) / (1 + s dt / 2e);  Cb = (dt / (e dx)) / (1 + s dt / 2e)
    with e = eps_r * eps0.  Raises RuntimeError when the permittivity and
    conductivity shapes differ (materials.py:47-49).
    """
    g = params.geom_params
    nx, ny, nz = g.size_x, g.size_y, g.size_z
    dx, dt = g.cell_size, g.timestep
    dtype = params.processing_params.datatype

    eps_r = materials.relative_permittivity
    sig = materials.conductivity
    if eps_r.shape != sig.shape:
        raise RuntimeError('Permittivity and conductivity arrays have '
                           'different shapes')
    n_vac = nz - g.size_z_material
    perm = np.concatenate((eps_r, np.ones((n_vac, ny, nx), dtype=dtype)),
                          axis=0, dtype=dtype)
    perm *= EPS0_2018
    cond = np.concatenate((sig, np.zeros((n_vac, ny, nx), dtype=dtype)),
                          axis=0, dtype=dtype)
    loss = (cond * dt) / (2.0 * perm)
    Ca = (1 - loss) / (1 + loss)
    Cb = (dt / (perm * dx)) / (1 + loss)
    assert Ca.dtype == dtype and Cb.dtype == dtype
    return Ca, Cb


def all_vacuum(params: ParameterCollection):
    """One vacuum model (materials.py:71-86)."""
    g = params.geom_params
    dtype = params.processing_params.datatype
    shape = (g.size_z_material, g.size_y, g.size_x)
    yield MaterialModel(np.ones(shape, dtype=dtype),
                        np.zeros(shape, dtype=dtype), g.cell_size, 'vacuum')


def flat_surface(params: ParameterCollection, epsilon_r: float,
                 conductivity: float, name='planar_surface'):
    """One homogeneous slab filling the material height (materials.py:89-109)."""
    g = params.geom_params
    dtype = params.processing_params.datatype
    shape = (g.size_z_material, g.size_y, g.size_x)
    meta = {'epsilon_r': epsilon_r, 'conductivity': conductivity}
    yield MaterialModel(epsilon_r * np.ones(shape, dtype=dtype),
                        conductivity * np.ones(shape, dtype=dtype),
                        g.cell_size, name, meta)


def fill_below_surface(eps, sigma, surf_height, epsilon_r, conductivity):
    """Set cells with z <= surf_height[y, x] to the given material, in place.

    The comparison is the reference's (materials.py:142-148,
    soilmodels.py:41-50): integer z against the height array as given
    (float heights for ``rough_dielectric``, rint'ed int32 for the soil
    generators).
    """
    nz = eps.shape[0]
    z = np.arange(nz).reshape(nz, 1, 1)
    mask = z <= surf_height[np.newaxis, :, :]
    eps[mask] = epsilon_r
    sigma[mask] = conductivity
    return mask


def rough_dielectric(params: ParameterCollection, epsilon_r: float,
                     conductivity: float, height_rms: float, mean_depth: float,
                     corr_len: float, count: int):
    """``count`` dielectric half-spaces under Gaussian rough surfaces
    (materials.py:112-164).  As in the reference, no seed is passed to the
    surface generator, so each call draws fresh surfaces."""
    g = params.geom_params
    nx, ny, nz = g.size_x, g.size_y, g.size_z_material
    dtype = params.processing_params.datatype
    dx = g.cell_size
    mean_depth_cells = mean_depth / dx
    if mean_depth > nz * dx:
        raise ValueError('Requested mean soil depth is greater than the max '
                         'height of the simulator space.')
    meta = {'epsilon_r': epsilon_r, 'conductivity': conductivity,
            'height_rms': height_rms, 'mean_depth': mean_depth,
            'mean_depth_cells': mean_depth_cells, 'corr_len': corr_len}
    for i in range(count):
        eps = np.ones((nz, ny, nx), dtype=dtype)
        sigma = np.zeros((nz, ny, nx), dtype=dtype)
        surf = gaussian_correlated_surface(nx, ny, dx, corr_len, corr_len,
                                           height_rms)
        surf /= dx
        surf += mean_depth_cells
        fill_below_surface(eps, sigma, surf, epsilon_r, conductivity)
        yield MaterialModel(eps, sigma, dx, f'rough_surf_{i}', dict(meta))
