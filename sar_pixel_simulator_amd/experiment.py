"""Experiment driver: backend selection and the model loop
(fdtdsarsim/experiment.py:19-97)."""
from typing import Iterable

from .incidentbeam import IncidentBeam
from .materials.materialmodel import MaterialModel
from .materials.materials import compute_material_arrays
from .parameters import ParameterCollection, set_geom_params
from .processresults import process_all_results
from .utils.log import defaultLogSetup, getLog

logger = getLog(__name__)


def get_fdtd_simulator(params: ParameterCollection):
    """Simulator class for ``processing_type`` (experiment.py:19-38).

    'hip' selects the MI355X backend.  'gpu' names the reference's CUDA
    backend, which cannot run on this hardware; it is served by the HIP
    backend so existing parameter files run unchanged.  The reference's numba
    'cpu' backend is not part of this package.
    """
    proc = params.processing_params.processing_type.lower()
    if proc in ('hip', 'gpu'):
        if proc == 'gpu':
            logger.info("processing_type 'gpu' is served by the HIP backend")
        from .hip.main_sim import MainSimHip
        return MainSimHip
    if proc == 'cpu':
        raise ValueError("processing_type 'cpu' (the reference's numba backend) is not "
                         "provided by this package; use 'hip'")
    raise ValueError('Invalid processing type')


def run_experiment(params: ParameterCollection, materials: Iterable[MaterialModel]):
    """Simulate every material model, then post-process all results
    (experiment.py:41-97).  When the geometry is not set, it is taken from
    the first model: shape (nz_mat, ny, nx) and its cell size.  A model may
    also be a ``materials.device.MaterialSpec`` (the geometry must then be
    set): its Ca, Cb are formed on the device."""
    from .materials.device import MaterialSpec
    defaultLogSetup()
    g = params.geom_params
    geom_specified = not (g.size_x == 0 and g.size_y == 0 and g.size_z == 0)
    sim = None
    if geom_specified:
        sim = get_fdtd_simulator(params)(params)
        IncidentBeam(params).setup()
    for material in materials:
        if isinstance(material, MaterialSpec):
            if sim is None:
                raise ValueError('a MaterialSpec needs the geometry set in the parameters')
            sim.run_spec(material)
            continue
        if sim is None:
            nz_mat, ny, nx = material.conductivity.shape
            dx = material.cell_size
            set_geom_params(params, nx * dx, ny * dx, nz_mat * dx, g.simulation_duration,
                            dx, 'metre')
            sim = get_fdtd_simulator(params)(params)
            IncidentBeam(params).setup()
        Ca, Cb = compute_material_arrays(params, material)
        sim.run(Ca, Cb, material.name, material.metadata)
    process_all_results(params.processing_params.project_dir,
                        params.display_params.debug_plots)
