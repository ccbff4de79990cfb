"""Simulation parameters: the same dataclasses, JSON keys and derived geometry
as the reference's parameters module, so a ``parameters.json`` written for
fdtdsarsim loads unchanged.

Reference: fdtdsarsim/parameters.py:16-145 (dataclasses), :159-239
(``load_from_json``), :242-305 (``set_geom_params``), :308-356 (file names).
"""
import dataclasses
import json
import os.path
from dataclasses import dataclass
from typing import List, Union

from . import constants
from .utils import log
from .utils.rounding import round_to_odd

#: Processing types accepted by ``get_fdtd_simulator``.  'hip' selects the
#: MI355X backend; 'cpu' and 'gpu' name the reference's own backends.
PROCESSING_TYPES = ('cpu', 'gpu', 'hip')


@dataclass
class ProcessingParams:
    """Where and how to run (parameters.py:16-27)."""

    processing_type: str = 'cpu'
    gpu_num: Union[int, List[int]] = 0
    threads_per_proc: int = 1
    datatype: str = 'float32'
    project_dir: str = ''
    #: Phasor accumulation on the HIP backend: 'compensated' keeps a
    #: float-float running sum per sample; 'reference' rounds the running sum
    #: to complex64 after every step exactly as the CPU path does
    #: (cpu/field_updates.py:385-400).
    dft_accumulation: str = 'compensated'

    def __str__(self):
        return _stringify_dataclass_members(self)


@dataclass
class IncBeamParams:
    """Incident beam definition (parameters.py:30-54)."""

    frequency: float = 0
    polarization: str = 0
    theta_angle: float = 0
    phi_angle: float = 0
    duration: float = 0
    illumination_width: float = 0
    generate_new: bool = False

    def __str__(self):
        return _stringify_dataclass_members(self)


@dataclass
class GeomParams:
    """Grid and timing (parameters.py:57-86)."""

    cell_size: float = 0
    simulation_duration: float = 0
    timestep: float = 0
    size_t: int = 0
    size_t_inc: int = 0
    size_t_pulse: int = 0
    size_x: int = 0
    size_y: int = 0
    size_z: int = 0
    size_pml: int = 0
    size_z_material: int = 0
    tfsf_z: int = 0
    courant: float = 0

    def __str__(self):
        return _stringify_dataclass_members(self)


@dataclass
class MaterialParams:
    """Material fill fraction (parameters.py:89-103)."""

    material_fraction: float = 0.7

    def __str__(self):
        return _stringify_dataclass_members(self)


@dataclass
class DisplayParams:
    """Display switches (parameters.py:106-122).  Plotting is not part of this
    package; the fields are kept so parameter files and result attributes
    keep their shape."""

    display_limit: float = 0.0
    display_step: int = 10
    inc_results: bool = False
    inc_sim_movie: bool = False
    debug_plots: bool = False
    main_sim_movie: bool = False
    material_plot: bool = False

    def __str__(self):
        return _stringify_dataclass_members(self)


@dataclass
class ParameterCollection:
    """All parameters of one experiment (parameters.py:125-145)."""

    processing_params: ProcessingParams = dataclasses.field(
        default_factory=ProcessingParams)
    inc_beam_params: IncBeamParams = dataclasses.field(
        default_factory=IncBeamParams)
    geom_params: GeomParams = dataclasses.field(default_factory=GeomParams)
    material_params: MaterialParams = dataclasses.field(
        default_factory=MaterialParams)
    display_params: DisplayParams = dataclasses.field(
        default_factory=DisplayParams)

    def __str__(self):
        return ''.join(f'{getattr(self, f.name)}\n'
                       for f in dataclasses.fields(self))


def _stringify_dataclass_members(obj):
    if not dataclasses.is_dataclass(obj):
        raise ValueError('Can only print a dataclass')
    lines = [f'{type(obj).__name__}:']
    lines += [f'   {f.name}: {getattr(obj, f.name)}'
              for f in dataclasses.fields(obj)]
    return '\n'.join(lines) + '\n'


_GEOM_KEYS = ('size_x', 'size_y', 'size_z_material', 'simulation_duration',
              'cell_size', 'unit')


def load_from_json(filename: str) -> ParameterCollection:
    """Read a parameters.json file (parameters.py:159-239).

    Required keys are the reference's; ``gpu_num`` is optional (default 0).
    When all of size_x, size_y, size_z_material, simulation_duration,
    cell_size and unit are present the geometry is derived immediately,
    otherwise it is left for ``run_experiment`` to take from the first
    material model.  Optional extension key: ``dft_accumulation``.
    """
    logger = log.getLog(__name__)
    logger.info(f'Loading parameters from the file {filename}')
    with open(filename, 'r') as fh:
        user = json.loads(fh.read())
    return params_from_dict(user)


def params_from_dict(user: dict) -> ParameterCollection:
    """Build a ParameterCollection from an already-parsed parameters dict
    (same keys and rules as ``load_from_json``)."""
    pc = ParameterCollection()
    proc = pc.processing_params
    inc = pc.inc_beam_params
    disp = pc.display_params

    proc.processing_type = user['processing_type']
    proc.gpu_num = user.get('gpu_num', 0)
    proc.threads_per_proc = user['threads_per_proc']
    proc.datatype = user['datatype']
    proc.project_dir = user['project_dir']
    if 'dft_accumulation' in user:
        proc.dft_accumulation = user['dft_accumulation']

    inc.frequency = user['frequency_hz']
    inc.polarization = user['polarization']
    inc.theta_angle = user['theta_angle_deg']
    inc.phi_angle = user['phi_angle_deg']
    inc.duration = user['incident_duration']
    inc.illumination_width = user['illumination_width']
    inc.generate_new = user['generate_new']

    pc.material_params.material_fraction = user['material_fraction']

    disp.display_limit = 0.5
    disp.display_step = 10
    disp.inc_results = user['disp_inc_results']
    disp.inc_sim_movie = user['disp_inc_sim_movie']
    disp.debug_plots = user['disp_debug']
    disp.main_sim_movie = user['disp_main_sim_movie']
    disp.material_plot = user['disp_materials']

    pc.geom_params.simulation_duration = user['simulation_duration']
    if all(k in user for k in _GEOM_KEYS):
        set_geom_params(pc, user['size_x'], user['size_y'],
                        user['size_z_material'], user['simulation_duration'],
                        user['cell_size'], user['unit'])
    return pc


def set_geom_params(pc: ParameterCollection, size_x: float, size_y: float,
                    size_z_material: float, size_t: float, cell_size: float,
                    unit: str, size_pml: int = 10):
    """Derive grid and timing from physical sizes (parameters.py:242-305).

    The arithmetic is kept operation for operation because the cell counts
    depend on it: cell counts are float floor-divisions of the sizes by
    ``cell_size`` *in the caller's unit* (3.0 // 0.05 == 59.0), lateral and
    vertical counts are rounded to odd, the Courant number is 0.5 and the
    CPML is 10 cells (``size_pml`` is an extension knob, default 10).
    """
    g = pc.geom_params
    g.courant = 0.5
    g.size_pml = int(size_pml)
    c0 = constants.C0
    lam = c0 / pc.inc_beam_params.frequency

    u = unit.lower()
    if u in ('metre', 'meter'):
        g.cell_size = cell_size
    elif u == 'wavelength':
        g.cell_size = lam * cell_size
    else:
        raise ValueError(f'Unknown value of unit: {unit}')

    g.size_x = round_to_odd(size_x // cell_size)
    g.size_y = round_to_odd(size_y // cell_size)
    g.size_z_material = int(size_z_material // cell_size)
    g.size_z = round_to_odd(g.size_z_material
                            / pc.material_params.material_fraction)
    g.tfsf_z = g.size_z_material + int(0.3 * lam / g.cell_size)

    g.timestep = g.cell_size * g.courant / c0
    dt = g.timestep
    freq = pc.inc_beam_params.frequency
    g.size_t_pulse = round(1.2 * pc.inc_beam_params.duration * lam
                           / (c0 * dt))
    g.size_t_inc = round(1.2 * g.size_t_pulse)
    g.size_t = round(size_t / dt / freq)
    if g.size_t_inc > g.size_t:
        log.getLog(__name__).warning(
            'Reducing the number of timesteps in the incident beam')
        g.size_t_inc = g.size_t


def inc_filename(pc: ParameterCollection) -> str:
    """Incident-beam cache file (parameters.py:308-324).  The name keeps the
    reference's key tuple; the container is ``.npz`` (HDF5 is not available
    to this package)."""
    g = pc.geom_params
    b = pc.inc_beam_params
    fname = (f'inc_nx{g.size_x}_ny{g.size_y}_nt{g.size_t_inc}'
             f'_theta{round(b.theta_angle)}_phi{round(b.phi_angle)}'
             f'_{b.polarization}.npz')
    return os.path.join(inc_file_dir(pc.processing_params), fname)


def results_filename(pc: ParameterCollection, material_name: str) -> str:
    """Results file of one material model (parameters.py:327-334)."""
    return os.path.join(results_file_dir(pc.processing_params),
                        material_name + '.npz')


def inc_file_dir(proc_params: ProcessingParams) -> str:
    d = os.path.join(proc_params.project_dir, 'incident')
    os.makedirs(d, exist_ok=True)
    return d


def results_file_dir(proc_params: ProcessingParams) -> str:
    d = os.path.join(proc_params.project_dir, 'results')
    os.makedirs(d, exist_ok=True)
    return d
