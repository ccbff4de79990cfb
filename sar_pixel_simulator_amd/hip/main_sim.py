"""HIP backend of the main simulation: the drop-in for MainSimCpu/MainSimGpu.

Same protocol as the reference backends (gpu/main_sim.py:42-449,
cpu/main_sim.py:22-307): ``MainSimHip(params, mpi=False)`` allocates the
device state once per geometry, ``run(Ca, Cb, material_name, material_meta)``
skips models whose results file exists, otherwise uploads Ca/Cb, runs
size_t steps with the incident beam resident in HBM, and writes the phasors.
"""
import json
import os
import time

import numpy as np

from .. import _native as native
from ..absorbers import Absorbers
from ..incidentbeam import IncidentBeam, params_device
from ..parameters import ParameterCollection, results_filename
from ..utils import log
from . import common


def params_attrs(params: ParameterCollection):
    """Every dataclass field as a flat attribute dict (gpu/main_sim.py:430-437)."""
    import dataclasses
    out = {}
    for grp in dataclasses.fields(params):
        obj = getattr(params, grp.name)
        for f in dataclasses.fields(obj):
            out[f.name] = getattr(obj, f.name)
    return out


def _jsonable(v):
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, np.ndarray):
        return v.tolist()
    return v


def save_phasor_file(fname, phasors, inc_phasors, params, nph, material_meta):
    """Results file: phasors, incident phasors and attributes
    (gpu/main_sim.py:387-449); refuses to overwrite."""
    if os.path.exists(fname):
        raise RuntimeError('Overwriting a previous simulation. This is probably bad!')
    arrays = {'phasors/' + n: a for n, a in phasors.items()}
    for n in native.INC_PHASOR_NAMES:
        arrays['phasors/' + n] = inc_phasors[n]
    attrs = {'phasors': {'nx_phasor': nph[0], 'ny_phasor': nph[1], 'nz_phasor': nph[2]},
             'parameters': {k: _jsonable(v) for k, v in params_attrs(params).items()},
             'material_metadata': {k: _jsonable(v) for k, v in (material_meta or {}).items()}}
    arrays['attrs'] = np.array(json.dumps(attrs))
    tmp = fname + '.tmp.npz'
    np.savez(tmp, **arrays)
    os.replace(tmp, fname)


class MainSimHip:
    logger = log.getLog(__name__)

    def __init__(self, params: ParameterCollection, mpi=False):
        common.require_float32(params)
        self.params = params
        self.mpi = mpi
        self.gpu_num = params_device(params)
        native.init(self.gpu_num)
        self.absorbers = Absorbers(params)
        g = params.geom_params
        self.nx_ph = g.size_x - 2 * g.size_pml
        self.ny_ph = g.size_y - 2 * g.size_pml
        self.nz_ph = g.size_z - 2 * g.size_pml
        self.geom = common.main_geom(params)
        self.sim = common.new_main_sim(params, self.absorbers)
        self._beam = None

    def _attach_beam(self):
        if self.geom.nt_inc == 0:
            return None
        inc = IncidentBeam(self.params)
        inc.setup()
        if self._beam is not inc.beam:
            self.sim.attach_beam(inc.beam)
            self._beam = inc.beam
        return inc

    def simulate(self, Ca, Cb):
        """Upload Ca/Cb, run size_t steps and leave the phasors on the device.
        Returns the IncidentBeam used."""
        shape = (self.geom.nz, self.geom.ny, self.geom.nx)
        Ca = np.asarray(Ca).astype(np.float32)
        Cb = np.asarray(Cb).astype(np.float32)
        if Ca.shape != shape or Cb.shape != shape:
            raise ValueError(f'material arrays must have shape {shape}, '
                             f'got {Ca.shape} and {Cb.shape}')
        if not (np.isfinite(Ca).all() and np.isfinite(Cb).all()):
            raise ValueError('material arrays hold non-finite values')
        inc = self._attach_beam()
        self.sim.reset()
        self.sim.set_materials(Ca, Cb)
        self.sim.run(self.geom.nt)
        self.sim.sync()
        return inc

    def simulate_spec(self, spec):
        """As ``simulate``, with Ca/Cb formed on the device from a
        ``materials.device.MaterialSpec``."""
        from ..materials import device
        inc = self._attach_beam()
        self.sim.reset()
        self.sim.build_materials(*device.to_c(spec, self.params))
        self.sim.run(self.geom.nt)
        self.sim.sync()
        return inc

    def run_spec(self, spec):
        """``run`` for a model described by a MaterialSpec."""
        t0 = time.monotonic()
        out_name = results_filename(self.params, spec.name)
        if os.path.exists(out_name):
            MainSimHip.logger.info(f'{out_name} already exists, skipping')
            return
        inc = self.simulate_spec(spec)
        self.save_results(spec.name, spec.metadata, inc)
        MainSimHip.logger.info(f'Simulation of {spec.name} on GPU {self.gpu_num} '
                               f'is complete ({time.monotonic() - t0:.3f}s)')

    def run(self, Ca: np.ndarray, Cb: np.ndarray, material_name: str, material_meta: dict):
        t0 = time.monotonic()
        out_name = results_filename(self.params, material_name)
        if os.path.exists(out_name):
            MainSimHip.logger.info(f'{out_name} already exists, skipping')
            return
        inc = self.simulate(Ca, Cb)
        self.save_results(material_name, material_meta, inc)
        MainSimHip.logger.info(f'Simulation of {material_name} on GPU {self.gpu_num} '
                               f'is complete ({time.monotonic() - t0:.3f}s)')

    def phasors(self):
        return self.sim.phasors()

    def far_field(self, thetas_deg, r=2.0e3):
        """Far field from the device-resident top-face phasors."""
        p = self.params.inc_beam_params
        return self.sim.near2far(thetas_deg, p.phi_angle, r, p.frequency)

    def save_results(self, material_name, material_meta, inc=None):
        if inc is None:
            inc = IncidentBeam(self.params)
            inc.setup()
        save_phasor_file(results_filename(self.params, material_name), self.sim.phasors(),
                         inc.phasors, self.params, (self.nx_ph, self.ny_ph, self.nz_ph),
                         material_meta)

    def close(self):
        self.sim.destroy()
