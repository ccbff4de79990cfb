"""Seeded material generators for the benchmark and parity workloads.

These are input producers, not part of the reference's API: soil with
per-cell random permittivity under a Gaussian rough surface, a voxelised
snow layer over that soil, and a vegetation canopy of random dielectric
cylinders.  The surface model and the fill rule (z <= rint(mean + h/dx))
follow materials/roughsurface.py:12-68 and materials/soilmodels.py:43-62.
"""
import numpy as np

from .materialmodel import MaterialModel
from .materials import fill_below_surface
from .roughsurface import gaussian_correlated_surface

#: Soil statistics shared by the workloads: eps_r ~ N(8, 0.5^2) clipped at 1,
#: sigma ~ U(0.005, 0.02) S/m, rough surface rms 0.02 lambda, Gaussian
#: correlation length 0.2 lambda.
SOIL = dict(eps_mean=8.0, eps_std=0.5, sig_lo=0.005, sig_hi=0.02,
            rms_wl=0.02, corr_wl=0.2)
#: Snow: binary ice/air voxels, ice volume fraction 0.3, eps_r 3.15,
#: sigma 1e-4 S/m.
SNOW = dict(ice_frac=0.3, ice_eps=3.15, ice_sig=1e-4)


def surface_cells(nx, ny, dx, lam, mean_cells, seed):
    """Interface height in cells, rint'ed to int32 (soilmodels.py:52-55)."""
    s = gaussian_correlated_surface(nx, ny, dx, SOIL['corr_wl'] * lam,
                                    SOIL['corr_wl'] * lam,
                                    SOIL['rms_wl'] * lam, seed=seed)
    return np.rint(s / dx + mean_cells).astype(np.int32)


def soil_fill(eps, sigma, height_cells, rng):
    """Fill z <= height with random soil drawn from ``rng``."""
    shape = eps.shape
    e = rng.normal(SOIL['eps_mean'], SOIL['eps_std'], size=shape)
    np.clip(e, 1.0, None, out=e)
    s = rng.uniform(SOIL['sig_lo'], SOIL['sig_hi'], size=shape)
    nz = shape[0]
    z = np.arange(nz).reshape(nz, 1, 1)
    mask = z <= height_cells[np.newaxis]
    eps[mask] = e[mask]
    sigma[mask] = s[mask]
    return mask


def rough_soil(nx, ny, nz_mat, dx, lam, mean_cells, seed, name=None,
               dtype='float32'):
    """Soil half-space under a rough surface, air above (the plain CPU and
    batched workloads)."""
    eps = np.ones((nz_mat, ny, nx), dtype=dtype)
    sig = np.zeros((nz_mat, ny, nx), dtype=dtype)
    h = surface_cells(nx, ny, dx, lam, mean_cells, seed)
    soil_fill(eps, sig, h, np.random.default_rng(seed + 1))
    meta = {'seed': seed, 'mean_cells': mean_cells}
    return MaterialModel(eps, sig, dx, name or f'soil_{seed}', meta)


def snow_over_soil(nx, ny, nz_mat, dx, lam, soil_cells, snow_cells, seed,
                   name=None, dtype='float32'):
    """Rough soil of mean depth ``soil_cells`` under a ``snow_cells`` thick
    layer of Bernoulli(0.3) ice voxels (the single-large-domain workload)."""
    eps = np.ones((nz_mat, ny, nx), dtype=dtype)
    sig = np.zeros((nz_mat, ny, nx), dtype=dtype)
    h = surface_cells(nx, ny, dx, lam, soil_cells, seed)
    rng = np.random.default_rng(seed + 2)
    top = min(nz_mat, soil_cells + snow_cells)
    ice = rng.random((top, ny, nx)) < SNOW['ice_frac']
    e_top = eps[:top]
    s_top = sig[:top]
    e_top[ice] = SNOW['ice_eps']
    s_top[ice] = SNOW['ice_sig']
    soil_fill(eps, sig, h, np.random.default_rng(seed + 1))
    meta = {'seed': seed, 'soil_cells': soil_cells, 'snow_cells': snow_cells}
    return MaterialModel(eps, sig, dx, name or f'snow_{seed}', meta)


def canopy_over_soil(nx, ny, nz_mat, dx, lam, soil_cells, seed, density,
                     name=None, dtype='float32'):
    """Random dielectric cylinders over rough soil (the serving workload).

    Cylinders: ``density`` per lambda^2 of footprint, radius U(0.05, 0.3)
    lambda, length U(0.5, 2) lambda, axis uniform on the sphere, eps_r
    U(10, 25), sigma U(0.1, 0.5) S/m; centres uniform over the footprint and
    between the soil surface and the top of the material volume.
    """
    eps = np.ones((nz_mat, ny, nx), dtype=dtype)
    sig = np.zeros((nz_mat, ny, nx), dtype=dtype)
    h = surface_cells(nx, ny, dx, lam, soil_cells, seed)
    rng = np.random.default_rng(seed + 3)
    area = (nx * dx / lam) * (ny * dx / lam)
    count = int(round(density * area))
    zc = np.arange(nz_mat) * dx
    yc = np.arange(ny) * dx
    xc = np.arange(nx) * dx
    for _ in range(count):
        r = rng.uniform(0.05, 0.3) * lam
        ln = rng.uniform(0.5, 2.0) * lam
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c = np.array([rng.uniform(soil_cells * dx, nz_mat * dx),
                      rng.uniform(0, ny * dx), rng.uniform(0, nx * dx)])
        e_c = rng.uniform(10.0, 25.0)
        s_c = rng.uniform(0.1, 0.5)
        # axis-aligned extent of the cylinder: L/2 |u_i| + r sqrt(1 - u_i^2)
        half = 0.5 * ln * np.abs(u) + r * np.sqrt(np.maximum(0.0, 1.0 - u * u))
        lo = np.maximum(np.floor((c - half) / dx).astype(int), 0)
        hi = np.minimum(np.ceil((c + half) / dx).astype(int) + 1,
                        [nz_mat, ny, nx])
        if np.any(hi <= lo):
            continue
        dz = (zc[lo[0]:hi[0]] - c[0])[:, None, None]
        dy = (yc[lo[1]:hi[1]] - c[1])[None, :, None]
        dxx = (xc[lo[2]:hi[2]] - c[2])[None, None, :]
        t = dz * u[0] + dy * u[1] + dxx * u[2]
        perp2 = dz * dz + dy * dy + dxx * dxx - t * t
        inside = (np.abs(t) <= 0.5 * ln) & (perp2 <= r * r)
        sub_e = eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        sub_s = sig[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        sub_e[inside] = e_c
        sub_s[inside] = s_c
    soil_fill(eps, sig, h, np.random.default_rng(seed + 1))
    meta = {'seed': seed, 'soil_cells': soil_cells, 'cylinders': count}
    return MaterialModel(eps, sig, dx, name or f'canopy_{seed}', meta)


__all__ = ['SOIL', 'SNOW', 'surface_cells', 'soil_fill', 'rough_soil',
           'snow_over_soil', 'canopy_over_soil', 'fill_below_surface']
