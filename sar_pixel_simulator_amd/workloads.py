"""Seeded workloads: parameter sets and material models for the plain,
single-large, batched and serving runs.

Geometry and timing come from ``set_geom_params`` or, for explicit cell
counts, from ``explicit_geometry`` which applies the same timing rules
(parameters.py:242-305: Courant 0.5, 10-cell CPML).
"""
import numpy as np

from .constants import C0
from .materials import synthetic
from .parameters import ParameterCollection, set_geom_params


def base_params(frequency=5.4e9, polarization='HH', theta=20.0, phi=0.0,
                incident_duration=10.0, illumination_width=2.0,
                simulation_duration=12.0, material_fraction=0.7,
                project_dir='', processing_type='hip', datatype='float32',
                dft_accumulation='compensated'):
    pc = ParameterCollection()
    p, b = pc.processing_params, pc.inc_beam_params
    p.processing_type, p.datatype, p.project_dir = processing_type, datatype, project_dir
    p.dft_accumulation = dft_accumulation
    b.frequency, b.polarization = frequency, polarization
    b.theta_angle, b.phi_angle = theta, phi
    b.duration, b.illumination_width = incident_duration, illumination_width
    pc.material_params.material_fraction = material_fraction
    pc.geom_params.simulation_duration = simulation_duration
    return pc


def explicit_geometry(pc: ParameterCollection, nx, ny, nz, nz_mat, cell_wl,
                      npml=10, tfsf_z=None):
    """Set the grid to explicit cell counts with the reference's timing rules."""
    g = pc.geom_params
    lam = C0 / pc.inc_beam_params.frequency
    g.courant, g.size_pml = 0.5, npml
    g.cell_size = lam * cell_wl
    g.size_x, g.size_y, g.size_z, g.size_z_material = nx, ny, nz, nz_mat
    g.tfsf_z = nz_mat + int(0.3 * lam / g.cell_size) if tfsf_z is None else tfsf_z
    g.timestep = g.cell_size * g.courant / C0
    dt, f = g.timestep, pc.inc_beam_params.frequency
    g.size_t_pulse = round(1.2 * pc.inc_beam_params.duration * lam / (C0 * dt))
    g.size_t_inc = round(1.2 * g.size_t_pulse)
    g.size_t = round(g.simulation_duration / dt / f)
    if g.size_t_inc > g.size_t:
        g.size_t_inc = g.size_t
    return pc


def lam_of(pc):
    return C0 / pc.inc_beam_params.frequency


# ---------------------------------------------------------------------------
# Plain CPU-sized run: 5.4 GHz HH, theta 20, 3x3 lambda lateral + 10-cell CPML
# (81 cells), 20 soil + 40 air cells + CPML in z (81), cell 0.05 lambda,
# rough soil interface at 30 cells, 480 steps.
# ---------------------------------------------------------------------------
def plain_params(**kw):
    pc = base_params(incident_duration=10.0, illumination_width=2.0,
                     simulation_duration=12.0, **kw)
    return explicit_geometry(pc, 81, 81, 81, 32, 0.05)


def plain_material(pc, seed=0):
    g = pc.geom_params
    return synthetic.rough_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                                lam_of(pc), 30, seed)


def small_params(nx=33, ny=31, nz=35, nz_mat=14, steps=None, **kw):
    """Reduced grid with the plain run's physics, for quick parity checks."""
    pc = base_params(incident_duration=2.0, illumination_width=1.0,
                     simulation_duration=3.0, **kw)
    explicit_geometry(pc, nx, ny, nz, nz_mat, 0.05)
    if steps is not None:
        g = pc.geom_params
        g.size_t = steps
        g.size_t_inc = min(g.size_t_inc, steps)
    return pc


# ---------------------------------------------------------------------------
# Single large run (README shape): 16x16x3 lambda at 0.021 lambda ->
# 761 x 761 x 203, 1714 steps; snow over soil.
# ---------------------------------------------------------------------------
def large_params(**kw):
    pc = base_params(incident_duration=10.0, illumination_width=12.0,
                     simulation_duration=18.0, material_fraction=0.7, **kw)
    set_geom_params(pc, 16.0, 16.0, 3.0, 18.0, 0.021, 'wavelength')
    return pc


def large_material(pc, seed=0):
    g = pc.geom_params
    nz_mat = g.size_z_material
    snow = int(round(1.0 / 0.021))
    soil = nz_mat - snow
    return synthetic.snow_over_soil(g.size_x, g.size_y, nz_mat, g.cell_size, lam_of(pc),
                                    soil, snow, seed)


# ---------------------------------------------------------------------------
# Largest single domain: 32 x 32 lambda lateral, 4 lambda material depth,
# 0.021 lambda cells, material fraction 0.61 (1523 x 1523 x 311 = 7.2e8
# cells), 21 periods (2000 steps), VV, theta 30; 1.5 lambda of snow over
# rough soil.  ``lateral`` shrinks the lateral size only (the same depth,
# cells, statistics and illumination as a fraction of the width): 8 lambda
# gives 381 x 381 x 311, small enough for the FP64 oracle.
# ---------------------------------------------------------------------------
def largest_params(lateral=32.0, **kw):
    pc = base_params(polarization='VV', theta=30.0, incident_duration=10.0,
                     illumination_width=0.75 * lateral, simulation_duration=21.0,
                     material_fraction=0.61, **kw)
    set_geom_params(pc, lateral, lateral, 4.0, 21.0, 0.021, 'wavelength')
    return pc


def largest_material(pc, seed=0):
    g = pc.geom_params
    nz_mat = g.size_z_material
    snow = int(round(1.5 / 0.021))
    return synthetic.snow_over_soil(g.size_x, g.size_y, nz_mat, g.cell_size, lam_of(pc),
                                    nz_mat - snow, snow, seed)


# ---------------------------------------------------------------------------
# Batched: 32 soil realizations, lateral {3,4,6,8} lambda, depth U[1,2]
# lambda, cell 0.03 lambda, HH and VV.
# ---------------------------------------------------------------------------
def batched_params_list(count=32, pols=('HH', 'VV'), seed=1000, datatype='float32'):
    rng = np.random.default_rng(seed)
    sizes = [3.0, 4.0, 6.0, 8.0]
    out = []
    for i in range(count):
        lat = sizes[i % 4]
        depth = float(rng.uniform(1.0, 2.0))
        for pol in pols:
            pc = base_params(polarization=pol, incident_duration=10.0,
                             illumination_width=2.0, simulation_duration=12.0,
                             material_fraction=0.5, datatype=datatype)
            set_geom_params(pc, lat, lat, depth, 12.0, 0.03, 'wavelength')
            out.append((i, pc))
    return out


def batched_material(pc, seed):
    g = pc.geom_params
    return synthetic.rough_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                                lam_of(pc), max(1, g.size_z_material - 4), seed)


def synthetic_beam_planes(pc, t0, n):
    """Analytic Gaussian-tapered plane wave at the TFSF plane, (4, n, ny, nx)
    float32 (ex, ey, hx, hy) for steps [t0, t0+n): a stand-in beam of the
    right shape for timing runs that do not generate the real one."""
    g, b = pc.geom_params, pc.inc_beam_params
    lam = lam_of(pc)
    k = 2 * np.pi / lam
    th = np.radians(b.theta_angle)
    x = (np.arange(g.size_x) - g.size_x / 2) * g.cell_size
    y = (np.arange(g.size_y) - g.size_y / 2) * g.cell_size
    w = b.illumination_width * lam / (2.0 * np.sqrt(2.0 * np.log(10.0)))
    env = np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / (w ** 2)).astype(np.float32)
    phase_x = (k * np.sin(th) * x)[None, :]
    out = np.empty((4, n, g.size_y, g.size_x), np.float32)
    om = 2 * np.pi * b.frequency
    for i in range(n):
        tt = (t0 + i) * g.timestep
        s = np.sin(om * tt - phase_x).astype(np.float32) * env
        out[0, i] = 0.0
        out[1, i] = 10.0 * s
        out[2, i] = -10.0 * s / 376.73
        out[3, i] = 0.0
    return out
