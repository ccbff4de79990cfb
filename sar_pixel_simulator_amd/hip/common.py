"""Scalars and launch descriptors shared by the HIP simulations."""
import numpy as np

from .. import _native as native
from ..absorbers import Absorbers
from ..constants import C0, EPS0_SCIPY, MU0_SCIPY
from ..parameters import ParameterCollection

DFT_MODES = {'compensated': native.DFT_COMPENSATED,
             'reference': native.DFT_REFERENCE}

#: Stepper of MAIN simulations (fdtd_sim_set_stepper): None = by grid size
#: (stepper_for), else 0 = split H and E kernels (in place), 1, 2, 3 = the
#: fused H+E kernels (double-buffered).  All give identical results.
FUSED_STEPPER = None

#: Grids of at least this many cells step fastest on the halo stepper (3),
#: smaller ones on the split stepper (0).  Measured on MI355X with the
#: default shapes: 45.1M cells split 0.704 vs halo 0.733 ms per step;
#: 85.1M 1.381 vs 1.336; 117.6M (761 x 761 x 203) 1.728 vs 1.600;
#: 721M (1523 x 1523 x 311) 10.21 vs 9.46; every batched soil model
#: (0.7-9M cells) runs 1.2-2x faster on the split stepper.
HALO_MIN_CELLS = 64_000_000


def stepper_for(nx, ny, nz):
    """Default stepper of a MAIN grid: 3 (halo) from HALO_MIN_CELLS cells
    up, else 0 (split)."""
    if FUSED_STEPPER is not None:
        return int(FUSED_STEPPER)
    return 3 if nx * ny * nz >= HALO_MIN_CELLS else 0


def new_main_sim(params, absorbers=None, fused=None):
    """A native.Sim for the main simulation of ``params`` with the default
    stepper for its grid (or ``fused``) and the tuned launch shape."""
    from . import tuning
    g = params.geom_params
    sim = native.Sim(main_geom(params), coeffs_for(params, absorbers))
    fused = stepper_for(g.size_x, g.size_y, g.size_z) if fused is None else fused
    sim.set_stepper(fused)
    shape = tuning.launch_for(g.size_x, g.size_y, g.size_z, fused=fused)
    try:
        sim.set_launch(*shape)
    except native.FdtdError:
        # a cached shape the library does not accept (a cache from another
        # build): the stepper's default shape, which every grid takes
        sim.set_launch(*tuning.DEFAULT[tuning.STEPPERS[int(fused)]])
    return sim


def require_float32(params: ParameterCollection):
    dt = params.processing_params.datatype
    if dt != 'float32':
        raise ValueError(f"the HIP backend computes in float32; datatype '{dt}' "
                         'is not supported')


def dft_mode(params: ParameterCollection):
    name = getattr(params.processing_params, 'dft_accumulation', 'compensated')
    try:
        return DFT_MODES[name]
    except KeyError:
        raise ValueError(f'unknown dft_accumulation {name!r}') from None


def update_scalars(params: ParameterCollection):
    """Db, dtm0, dx, dte0dx as float32 and omega, dt as float64, formed as
    cpu/main_sim.py:101-122 forms them (scipy.constants eps0 / mu0)."""
    g = params.geom_params
    dx, dt = g.cell_size, g.timestep
    return dict(Db=np.float32(dt / (MU0_SCIPY * dx)),
                dtm0=np.float32(dt / MU0_SCIPY),
                dx=np.float32(dx),
                dte0dx=np.float32(dt / (EPS0_SCIPY * dx)),
                omega=2.0 * np.pi * params.inc_beam_params.frequency,
                dt=dt)


def coeffs_for(params: ParameterCollection, absorbers: Absorbers = None):
    ab = absorbers or Absorbers(params)
    s = update_scalars(params)
    return native.make_coeffs(s['Db'], s['dtm0'], s['dx'], s['dte0dx'],
                              s['omega'], s['dt'], ab.profiles(),
                              cell_size=params.geom_params.cell_size)


def main_geom(params: ParameterCollection):
    g = params.geom_params
    return native.make_geom(g.size_x, g.size_y, g.size_z, g.size_pml, g.tfsf_z,
                            g.size_t, g.size_t_inc, mode=native.MODE_MAIN,
                            dft_accum=dft_mode(params))


def incident_planes(params: ParameterCollection):
    """Height, feed-in and read-out planes of the incident simulation
    (cpu/incident_sim.py:34-37, :100-103): nz = int(3 lambda / dx),
    feed_in = nz // 2 + 4, read_out = feed_in - 4."""
    lam = C0 / params.inc_beam_params.frequency
    nz = int(3.0 * lam / params.geom_params.cell_size)
    offset = 8
    feed_in = (nz // 2) + (offset // 2)
    read_out = feed_in - (offset // 2)
    return nz, feed_in, read_out


def incident_geom(params: ParameterCollection):
    g = params.geom_params
    nz, feed_in, read_out = incident_planes(params)
    n_src = min(g.size_t_pulse, g.size_t_inc)
    return native.make_geom(g.size_x, g.size_y, nz, g.size_pml, feed_in,
                            g.size_t_inc, n_src, mode=native.MODE_INCIDENT,
                            read_out=read_out, dft_accum=dft_mode(params))


def incident_cb(params: ParameterCollection):
    """Cb of the incident simulation's vacuum, float32(dt / (eps0 dx)) with
    scipy's eps0 (cpu/incident_sim.py:116)."""
    g = params.geom_params
    return np.float32(g.timestep / (EPS0_SCIPY * g.cell_size))
