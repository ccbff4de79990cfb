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

    def lookup(self, params):
        """(beam, incident phasors) if resident, else None; no reference."""
        hit = self._beams.get(beam_key(params))
        return None if hit is None else hit[:2]

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


class SimPool:
    """Device simulations of finished models, kept for the next model on the
    same grid (the reference's backends likewise allocate once per geometry
    and zero the state per model, gpu/main_sim.py:102, :691-787).  A
    recycled simulation is reset (fields, CPML state, phasor sums, step
    counter) and takes the new model's beam and materials, so its results
    equal a fresh one's.  Freeing device memory waits for all work on the
    device; recycling avoids that between models.  ``evict`` frees the least
    recently kept ones when memory is needed."""

    def __init__(self):
        self._free = {}       # key -> [(tick, sim, bytes)]
        self._tick = 0

    @staticmethod
    def key(params):
        return (bytes(common.main_geom(params)), bytes(common.coeffs_for(params)))

    def has(self, params):
        return bool(self._free.get(self.key(params)))

    def take(self, params):
        lst = self._free.get(self.key(params))
        if not lst:
            return None
        return lst.pop()[1]

    def give(self, params, sim, nbytes):
        self._tick += 1
        self._free.setdefault(self.key(params), []).append((self._tick, sim, nbytes))

    def bytes(self):
        return sum(nb for lst in self._free.values() for _, _, nb in lst)

    def evict(self, nbytes, keep=None):
        """Destroy kept simulations, least recently kept first, until at
        least ``nbytes`` are freed; never one for ``keep`` (params)."""
        keep_k = self.key(keep) if keep is not None else None
        entries = sorted((t, k, i) for k, lst in self._free.items() if k != keep_k
                         for i, (t, _, _) in enumerate(lst))
        freed, drop = 0, []
        for t, k, i in entries:
            if freed >= nbytes:
                break
            drop.append((k, t))
            freed += [nb for tt, _, nb in self._free[k] if tt == t][0]
        for k, t in drop:
            lst = self._free[k]
            for j, (tt, sim, _) in enumerate(lst):
                if tt == t:
                    sim.destroy()
                    del lst[j]
                    break
        return freed

    def release(self):
        for lst in self._free.values():
            for _, sim, _ in lst:
                sim.destroy()
        self._free.clear()


class DeviceModel:
    """One material model's simulation resident on the device."""

    def __init__(self, params, Ca, Cb, beams: BeamCache, name='model', meta=None, spec=None,
                 pool: SimPool = None):
        common.require_float32(params)
        self.params = params
        self.name = name
        self.meta = meta or {}
        g = params.geom_params
        self.nt = g.size_t
        self.cells = g.size_x * g.size_y * g.size_z
        self.geom = common.main_geom(params)
        self._pool = pool
        self.sim = pool.take(params) if pool is not None else None
        if self.sim is not None:
            self.sim.reset()
        else:
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
    def from_material(cls, params, material, beams, name=None, pool=None):
        Ca, Cb = compute_material_arrays(params, material)
        return cls(params, Ca, Cb, beams, name or material.name, material.metadata, pool=pool)

    @classmethod
    def from_spec(cls, params, spec, beams, name=None, pool=None):
        """A model whose Ca, Cb are formed on the device from a
        ``materials.device.MaterialSpec``."""
        return cls(params, None, None, beams, name or spec.name, spec.metadata, spec=spec,
                   pool=pool)

    @staticmethod
    def footprint(params):
        return native.footprint(common.main_geom(params))

    @property
    def done(self):
        return self.steps_done >= self.nt

    def far_field(self, thetas=THETAS, r=RADIUS):
        """Far field from the device-resident top-face phasors; through the
        entry point that allocates nothing when the sizes allow, so that
        collecting a finished member does not wait for a batch still
        running on the device."""
        b, g = self.params.inc_beam_params, self.params.geom_params
        side = max(g.size_x, g.size_y) - 2 * g.size_pml
        noalloc = np.size(thetas) <= native.N2F_MAX_THETA and side <= native.N2F_MAX_SIDE
        return self.sim.near2far(thetas, b.phi_angle, r, b.frequency, noalloc=noalloc)

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
        """Free the model's device state, or keep its simulation in the pool
        it came with for the next model on the same grid."""
        if self._pool is not None:
            self._pool.give(self.params, self.sim, self.footprint(self.params))
        else:
            self.sim.destroy()
        self.sim = None
        if self.beam is not None and self._beams is not None:
            self._beams.unref(self.params)
            self.beam = None


#: Members below this many cells are latency-bound when stepped alone (a
#: launch covers too few tiles to fill the device); measured on MI355X the
#: 0.7e6 to 9e6-cell soil models step 1.12x faster as four concurrent lanes
#: than one after another, while the 3e6 to 6e7-cell canopy models step
#: fastest one after another (concurrent lanes 3 % slower, one packed tile
#: list 6 to 10 % slower).
SMALL_CELLS = 8_000_000
SMALL_LANES = 4


def auto_mode(models):
    """(mode, small_cells, lanes) for a batch of ``models``: every member
    with its own launches, a chunk of steps per member; the small members
    on four concurrent lanes, then the large ones one after another."""
    small = sum(m.cells < SMALL_CELLS for m in models)
    if small == 0:
        return native.BATCH_SERIAL, 0, 1
    lanes = min(SMALL_LANES, small)
    if small == len(models):
        return native.BATCH_SERIAL, 0, lanes
    return native.BATCH_SERIAL, SMALL_CELLS, lanes


class DeviceBatch:
    """Members stepped together; rebuild when membership changes.

    ``mode`` is 'auto' (``auto_mode``), 'packed' (one tile list, one launch
    per kernel per step for all members, ``launch`` its tile shape) or a
    (mode, small_cells, lanes) tuple of fdtd_batch_set_mode.  Results are
    identical in every mode."""

    def __init__(self, models, launch=None, mode='auto'):
        self.models = list(models)
        self.batch = native.Batch([m.sim for m in self.models])
        if mode == 'auto':
            mode = auto_mode(self.models)
        if mode == 'packed':
            mode = (native.BATCH_PACKED, 0, 1)
            launch = launch or PACKED_LAUNCH
        self.mode = tuple(mode)
        if launch:
            self.batch.set_launch(*launch)
        if self.mode[0] != native.BATCH_PACKED:
            self.batch.set_mode(*self.mode)

    def run(self, nsteps, use_graph=None):
        """Queue nsteps steps of every member.  By default launches go to the
        streams directly (measured faster than graph replay for the serial
        modes) and the packed mode replays graphs."""
        if use_graph is None:
            use_graph = self.mode[0] == native.BATCH_PACKED
        self.batch.run(nsteps, use_graph)
        for m in self.models:
            m.steps_done = min(m.nt, m.steps_done + nsteps)

    def sync(self):
        self.batch.sync()

    def busy(self):
        return self.batch.busy()

    def destroy(self):
        self.batch.destroy()


#: Tile shape of the packed mode: each member its own block width, 4-plane z
#: chunks (measured best for the 64 soil models: 4.65 ms per step against
#: 4.88 with 2-plane chunks).
PACKED_LAUNCH = (0, 4, 0, 0, 0)


def run_plain(models, sync=True):
    """One model at a time, as the reference's experiment loop does."""
    for m in models:
        m.sim.run(m.nt)
        m.steps_done = m.nt
        if sync:
            m.sim.sync()


def run_batched(models, chunk=64, use_graph=None, launch=None, mode='auto'):
    """All models together; steps until the longest is done.  ``mode``
    'tuned' takes the execution the tuner cached for this batch's grids,
    measuring the candidates first when none is cached ('tune': always);
    otherwise as DeviceBatch."""
    if mode in ('tune', 'tuned'):
        from . import tuning
        hit = tuning.batch_mode_for(models) if mode == 'tuned' else None
        if hit is None:
            hit, _ = tuning.autotune_batch(models)
        mode, launch = hit
    b = DeviceBatch(models, launch, mode)
    nmax = max(m.nt for m in models)
    done = 0
    while done < nmax:
        n = min(chunk, nmax - done)
        b.run(n, use_graph)
        done += n
    b.sync()
    b.destroy()
