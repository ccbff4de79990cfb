"""Device-resident simulations and batched execution.

The reference runs one material model at a time per process and farms
models over MPI ranks (experiment.py:69-97, experiment_mpi.py:86-255).  Here
many models -- of different grid sizes, polarisations and angles -- live in
HBM together and advance in lock-step: one launch per kernel per time step
covers every member (libfdtd_hip batches), and the step sequence replays as
a captured HIP graph.  Each member keeps its own step counter, fields, CPML
state and phasor sums, so results are identical to running it alone.
"""
import numpy as np

from .. import _native as native
from ..incidentbeam import synthesize_on_device
from ..materials import compute_material_arrays
from ..processresults import THETAS, RADIUS, incident_power, scattering_coefficients, \
    farfield_power, backscatter
from . import common


class BeamCache:
    """Incident beams resident in HBM, keyed by what the beam depends on (the
    reference's incident-file key, parameters.py:308-324, plus the values its
    metadata check compares, incidentbeam.py:105-165).  A beam is generated on
    first use, entirely on the device (excitation synthesis and incident
    FDTD).  Each model holding a beam counts as a reference; a beam no model
    references stays cached for later models with the same key until
    ``evict`` frees it (least recently used first)."""

    def __init__(self):
        self._beams = {}      # key -> (beam, phasors, p_inc, bytes)
        self._refs = {}       # key -> models holding the beam
        self._used = {}       # key -> last use (a counter)
        self._tick = 0

    @staticmethod
    def beam_bytes(params):
        g = params.geom_params
        return 4 * 4 * g.size_t_inc * g.size_y * g.size_x

    def bytes_for(self, params):
        """HBM a model of ``params`` adds for its beam: 0 when resident."""
        if params.geom_params.size_t_inc == 0 or beam_key(params) in self._beams:
            return 0
        return self.beam_bytes(params)

    def resident_bytes(self):
        return sum(v[3] for v in self._beams.values())

    def get(self, params):
        """(beam, incident phasors, incident power), generated if absent;
        counts a reference until ``unref``."""
        k = beam_key(params)
        hit = self._beams.get(k)
        if hit is None:
            beam, ph = synthesize_on_device(params)
            hit = (beam, ph, incident_power(ph, params.geom_params.cell_size),
                   self.beam_bytes(params))
            self._beams[k] = hit
        self._refs[k] = self._refs.get(k, 0) + 1
        self._tick += 1
        self._used[k] = self._tick
        return hit[:3]

    def warm(self, params):
        """Generate the beam of ``params`` ahead of use, holding no reference."""
        self.get(params)
        self.unref(params)

    def unref(self, params):
        k = beam_key(params)
        if self._refs.get(k, 0) > 0:
            self._refs[k] -= 1

    def put(self, params, beam, inc_phasors):
        k = beam_key(params)
        self._beams[k] = (beam, inc_phasors,
                          incident_power(inc_phasors, params.geom_params.cell_size),
                          self.beam_bytes(params))
        self._refs.setdefault(k, 0)
        self._tick += 1
        self._used[k] = self._tick

    def evict(self, nbytes, keep=None):
        """Free unreferenced beams, least recently used first, until at least
        ``nbytes`` are freed; never the beam of ``keep`` (params).  Returns
        the bytes freed."""
        keep_k = beam_key(keep) if keep is not None else None
        freed = 0
        for k in sorted(self._beams, key=lambda k: self._used.get(k, 0)):
            if freed >= nbytes:
                break
            if k == keep_k or self._refs.get(k, 0) > 0:
                continue
            b, _, _, nb = self._beams.pop(k)
            b.release()
            self._refs.pop(k, None)
            self._used.pop(k, None)
            freed += nb
        return freed

    def __len__(self):
        return len(self._beams)

    def release(self):
        for b, _, _, _ in self._beams.values():
            b.release()
        self._beams.clear()
        self._refs.clear()
        self._used.clear()


def beam_key(params):
    g, b = params.geom_params, params.inc_beam_params
    return (g.size_x, g.size_y, g.size_t_inc, round(b.theta_angle), round(b.phi_angle),
            b.polarization, g.cell_size, b.frequency, b.duration, b.illumination_width,
            params.processing_params.dft_accumulation)


class DeviceModel:
    """One material model's simulation resident on the device."""

    def __init__(self, params, Ca, Cb, beams: BeamCache, name='model', meta=None, spec=None):
        common.require_float32(params)
        self.params = params
        self.name = name
        self.meta = meta or {}
        g = params.geom_params
        self.nt = g.size_t
        self.geom = common.main_geom(params)
        self.sim = common.new_main_sim(params)
        self.beam, self.inc_phasors, self.p_inc = (None, None, None)
        self._beams = beams
        if g.size_t_inc > 0:
            self.beam, self.inc_phasors, self.p_inc = beams.get(params)
            self.sim.attach_beam(self.beam)
        if spec is not None:
            from ..materials import device
            self.sim.build_materials(*device.to_c(spec, params))
        else:
            self.sim.set_materials(Ca, Cb)
        self.steps_done = 0

    @classmethod
    def from_material(cls, params, material, beams, name=None):
        Ca, Cb = compute_material_arrays(params, material)
        return cls(params, Ca, Cb, beams, name or material.name, material.metadata)

    @classmethod
    def from_spec(cls, params, spec, beams, name=None):
        """A model whose Ca, Cb are formed on the device from a
        ``materials.device.MaterialSpec``."""
        return cls(params, None, None, beams, name or spec.name, spec.metadata, spec=spec)

    @staticmethod
    def footprint(params):
        return native.footprint(common.main_geom(params))

    @property
    def done(self):
        return self.steps_done >= self.nt

    def far_field(self, thetas=THETAS, r=RADIUS):
        b = self.params.inc_beam_params
        return self.sim.near2far(thetas, b.phi_angle, r, b.frequency)

    def sigma0(self, thetas=THETAS, r=RADIUS):
        """Far field and (sigma_h, sigma_v) over ``thetas`` (processresults.py:
        104-121), plus the backscatter values at theta = -theta_inc."""
        g = self.params.geom_params
        ff = self.far_field(thetas, r)
        ph, pv = farfield_power(ff['E_theta'], ff['E_phi'], ff['H_theta'], ff['H_phi'])
        n = g.size_pml
        sh, sv = scattering_coefficients(ph, pv, self.p_inc, g.size_x - 2 * n,
                                         g.size_y - 2 * n, g.cell_size, r)
        back = None
        th_inc = self.params.inc_beam_params.theta_angle
        if np.any(np.asarray(thetas) == -th_inc):
            back = backscatter(sh, sv, th_inc, thetas)
        return {'far_field': ff, 'sigma_h': sh, 'sigma_v': sv, 'backscatter': back}

    def release(self):
        self.sim.destroy()
        if self.beam is not None and self._beams is not None:
            self._beams.unref(self.params)
            self.beam = None


class DeviceBatch:
    """Members stepped together; rebuild when membership changes."""

    def __init__(self, models, launch=None):
        self.models = list(models)
        self.batch = native.Batch([m.sim for m in self.models])
        if launch:
            self.batch.set_launch(*launch)

    def run(self, nsteps, use_graph=True):
        self.batch.run(nsteps, use_graph)
        for m in self.models:
            m.steps_done = min(m.nt, m.steps_done + nsteps)

    def sync(self):
        self.batch.sync()

    def destroy(self):
        self.batch.destroy()


def run_plain(models, sync=True):
    """One model at a time, as the reference's experiment loop does."""
    for m in models:
        m.sim.run(m.nt)
        m.steps_done = m.nt
        if sync:
            m.sim.sync()


def run_batched(models, chunk=64, use_graph=True, launch=None):
    """All models together; steps until the longest is done."""
    b = DeviceBatch(models, launch)
    nmax = max(m.nt for m in models)
    done = 0
    while done < nmax:
        n = min(chunk, nmax - done)
        b.run(n, use_graph)
        done += n
    b.sync()
    b.destroy()
