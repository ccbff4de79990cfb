"""Shared set-up for the parity tests: the same seeded inputs for the HIP
library and the oracle."""
import numpy as np

import beam_oracle
import oracle
from sar_pixel_simulator_amd import workloads
from sar_pixel_simulator_amd.absorbers import Absorbers
from sar_pixel_simulator_amd.hip import common
from sar_pixel_simulator_amd.materials import compute_material_arrays


def profiles_of(params):
    ft = params.processing_params.datatype
    return np.concatenate([np.asarray(p, ft) for p in Absorbers(params).profiles()])


def scalars(params):
    """Update scalars in the configured float type (float32 for the HIP
    library, float64 for the FP64 oracle)."""
    s = common.update_scalars(params)
    if params.processing_params.datatype == 'float64':
        from sar_pixel_simulator_amd.constants import EPS0_SCIPY, MU0_SCIPY
        g = params.geom_params
        dx, dt = g.cell_size, g.timestep
        s.update(Db=np.float64(dt / (MU0_SCIPY * dx)), dtm0=np.float64(dt / MU0_SCIPY),
                 dx=np.float64(dx), dte0dx=np.float64(dt / (EPS0_SCIPY * dx)))
    return s


def oracle_geom(params):
    g = params.geom_params
    s = scalars(params)
    return dict(nx=g.size_x, ny=g.size_y, nz=g.size_z, npml=g.size_pml,
                tfsf_z=g.tfsf_z, nt_inc=g.size_t_inc, Db=s['Db'], dtm0=s['dtm0'],
                dx=s['dx'], dte0dx=s['dte0dx'], omega=s['omega'], dt=s['dt'])


def oracle_incident_geom(params):
    g = params.geom_params
    s = scalars(params)
    nz, feed_in, read_out = common.incident_planes(params)
    return dict(nx=g.size_x, ny=g.size_y, nz=nz, npml=g.size_pml, tfsf_z=feed_in,
                read_out=read_out, nt_inc=min(g.size_t_pulse, g.size_t_inc),
                Db=s['Db'], dtm0=s['dtm0'], dx=s['dx'], dte0dx=0.0,
                omega=s['omega'], dt=s['dt'])


def soil_case(nx=33, ny=31, nz=35, nz_mat=14, steps=None, seed=3, **kw):
    """Small rough-soil case: params, Ca, Cb."""
    pc = workloads.small_params(nx=nx, ny=ny, nz=nz, nz_mat=nz_mat, steps=steps, **kw)
    g = pc.geom_params
    from sar_pixel_simulator_amd.materials import synthetic
    m = synthetic.rough_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                             workloads.lam_of(pc), max(1, nz_mat - 4), seed)
    Ca, Cb = compute_material_arrays(pc, m)
    return pc, Ca, Cb


def synthetic_beam(pc):
    g = pc.geom_params
    return workloads.synthetic_beam_planes(pc, 0, g.size_t_inc)


def excitation(pc):
    return beam_oracle.create_inc_excitation(pc)


def oracle_incident(pc, pulse_ex, pulse_ey, nthreads=None):
    """Oracle run of the incident simulation: (beam (4, nt, ny, nx), phasors)."""
    g = pc.geom_params
    ft = pc.processing_params.datatype
    geo = oracle_incident_geom(pc)
    ny, nx = g.size_y, g.size_x
    nz = geo['nz']
    n_src = geo['nt_inc']
    src = np.stack([np.imag(pulse_ex[:n_src]), np.imag(pulse_ey[:n_src])]).astype(ft)
    ones = np.ones((nz, ny, nx), ft)
    from sar_pixel_simulator_amd.constants import EPS0_SCIPY
    cb_val = common.incident_cb(pc) if ft == 'float32' else \
        g.timestep / (EPS0_SCIPY * g.cell_size)
    cb = np.full((nz, ny, nx), cb_val, ft)
    sim = oracle.OracleSim(geo, ones, cb, profiles_of(pc), inc=src, nthreads=nthreads,
                           incident=True, dtype=ft)
    rec = sim.run_incident(g.size_t_inc)
    return rec, sim.incident_phasors()


def oracle_main(pc, Ca, Cb, beam, steps=None, nthreads=None):
    g = pc.geom_params
    sim = oracle.OracleSim(oracle_geom(pc), Ca, Cb, profiles_of(pc), inc=beam,
                           nthreads=nthreads, dtype=pc.processing_params.datatype)
    sim.run(g.size_t if steps is None else steps)
    return sim


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)
