"""Single-GPU serving loop with continuous batching.

Replaces the reference's one-process-per-GPU MPI farm
(experiment_mpi.py:86-255: rank 0 sends Ca/Cb to an idle worker, the worker
runs one model and reports SUCCESS).  Here requests queue on the host;
at every chunk boundary the loop

1. waits for the chunk in flight and frees the members retired at the
   previous boundary;
2. takes finished (and cancelled) members out of the batch;
3. rebuilds the batch when membership changed and starts the next chunk of
   time steps for every member (each member's chunk with its own launches,
   one member after another or on a few concurrent lanes, or all members in
   one tile list: hip.batch.DeviceBatch);
4. while that chunk runs on the device: admits queued requests in arrival
   order while a slot is free and the request's HBM footprint (fields,
   coefficients, CPML state, phasor sums, plus its incident beam unless
   already resident) fits the budget -- they join at the next boundary --,
   prepares queued requests on the host (the Yee coefficients of requests
   that carry a MaterialModel) and collects the retired members' far field
   and sigma0 from their device-resident phasor sums.

Requests that arrive while others run join at a boundary, so their steps
overlap on the device.  The HBM in use is the members' own state plus
every incident beam resident on the device: a beam is counted once, from
its creation until it is freed, whichever requests share it; beams no
member references are kept for later requests and evicted (least recently
used first) when an admission needs their memory; so are the device
simulations of finished requests, kept for the next request on the same
grid (hip.batch.SimPool).  ``submit`` refuses a
request whose inputs are malformed (shape, non-finite or non-positive
values) or that could never fit; ``cancel`` withdraws a queued request at
once and a running one at the next boundary.  Each member's arithmetic
does not depend on the other members, so a result is bit-identical alone
and in the mix.

The device side is behind a small backend interface (``HipBackend``) so the
scheduling logic can be exercised without a GPU.
"""
import collections
import dataclasses
import time
from typing import Any, Callable, Optional

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    params: Any                      # ParameterCollection with geometry set
    material: Any = None             # MaterialModel, or None when Ca/Cb or spec given
    spec: Any = None                 # materials.device.MaterialSpec: built on the device
    Ca: Optional[np.ndarray] = None
    Cb: Optional[np.ndarray] = None
    name: str = ''
    arrival: float = 0.0             # seconds on the server clock
    prepared: bool = False           # Ca, Cb computed by the backend while queued


@dataclasses.dataclass
class Result:
    rid: int
    name: str
    arrival: float
    admitted: float
    finished: float
    steps: int
    output: Any = None
    error: Optional[str] = None

    @property
    def latency(self):
        return self.finished - self.arrival


class AdmissionError(RuntimeError):
    """The request can never fit the HBM budget."""


def _finite_positive(name, a, positive):
    a = np.asarray(a)
    if not np.isfinite(a).all():
        raise ValueError(f'{name} holds non-finite values')
    if positive and not (a > 0).all():
        raise ValueError(f'{name} must be positive')


def check_request(req):
    """ValueError unless the request's material inputs have the shapes its
    geometry gives and finite values: Ca, Cb (nz, ny, nx); a MaterialModel's
    permittivity (> 0) and conductivity (nz_mat, ny, nx), as
    compute_material_arrays expects (materials/materials.py:23-68).  A
    MaterialSpec is checked by the device builder."""
    g = req.params.geom_params
    if req.spec is not None:
        return
    if req.Ca is not None or req.Cb is not None:
        want = (g.size_z, g.size_y, g.size_x)
        for name, a in (('Ca', req.Ca), ('Cb', req.Cb)):
            if a is None or np.shape(a) != want:
                raise ValueError(f'{name} must have shape {want}, got '
                                 f'{None if a is None else np.shape(a)}')
            _finite_positive(name, a, False)
        return
    m = req.material
    if m is None:
        raise ValueError(f'request {req.rid} carries no material')
    want = (g.size_z_material, g.size_y, g.size_x)
    for name, a, pos in (('relative_permittivity', m.relative_permittivity, True),
                         ('conductivity', m.conductivity, False)):
        if np.shape(a) != want:
            raise ValueError(f'{name} must have shape {want}, got {np.shape(a)}')
        _finite_positive(name, a, pos)


class HipBackend:
    """Device operations the loop needs, on libfdtd_hip."""

    def __init__(self, device=0, launch=None, use_graph=None, mode='auto', recycle=True):
        from .. import _native as native
        from .batch import BeamCache, SimPool
        native.init(device)
        self.native = native
        self.beams = BeamCache()
        self.pool = SimPool() if recycle else None
        self.launch = launch
        self.use_graph = use_graph          # None: the batch mode's own choice
        self.mode = mode

    def model_bytes(self, req):
        """Device state of the request's simulation (beam excluded)."""
        from .batch import DeviceModel
        return DeviceModel.footprint(req.params)

    def footprint(self, req):
        """Device memory an admission adds: the simulation unless a kept one
        on the same grid is recycled, and the beam unless resident."""
        recycled = self.pool is not None and self.pool.has(req.params)
        return (0 if recycled else self.model_bytes(req)) + self.beams.bytes_for(req.params)

    def cached_bytes(self):
        """Resident beams and kept simulations, whether or not in use."""
        return self.beams.resident_bytes() + (self.pool.bytes() if self.pool is not None else 0)

    def evict_cached(self, nbytes, keep):
        """Free kept simulations, then unreferenced beams, until ``nbytes``
        are freed; never what ``keep`` (a request) would reuse."""
        freed = self.pool.evict(nbytes, keep=keep.params) if self.pool is not None else 0
        if freed < nbytes:
            freed += self.beams.evict(nbytes - freed, keep=keep.params)
        return freed

    def check(self, req):
        check_request(req)

    def steps(self, req):
        return req.params.geom_params.size_t

    def prepare(self, req):
        """Host-side part of admit() that needs no device memory: the Yee
        coefficients of a request given as a MaterialModel."""
        from ..materials import compute_material_arrays
        if req.spec is None and req.Ca is None and req.material is not None:
            req.Ca, req.Cb = compute_material_arrays(req.params, req.material)
            req.prepared = True

    def admit(self, req):
        from .batch import DeviceModel
        from ..materials import compute_material_arrays
        if req.spec is not None:
            return DeviceModel.from_spec(req.params, req.spec, self.beams, req.name, pool=self.pool)
        Ca, Cb = (req.Ca, req.Cb) if req.Ca is not None else \
            compute_material_arrays(req.params, req.material)
        model = DeviceModel(req.params, Ca, Cb, self.beams, req.name, pool=self.pool)
        if req.prepared:                    # uploaded: the host copy is no longer needed
            req.Ca = req.Cb = None
            req.prepared = False
        return model

    def make_batch(self, members):
        from .batch import DeviceBatch
        return DeviceBatch(members, self.launch, self.mode)

    def run(self, batch, nsteps):
        batch.run(nsteps, use_graph=self.use_graph)

    def sync(self, batch):
        batch.sync()

    def busy(self, batch):
        return batch.busy()

    def drop_batch(self, batch):
        batch.destroy()

    def collect(self, member):
        return member.sigma0()

    def release(self, member):
        member.release()


class Server:
    """Continuous-batching request loop on one device."""

    def __init__(self, backend, hbm_budget, max_slots=64, chunk_steps=32,
                 clock: Callable[[], float] = None):
        self.backend = backend
        self.budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.chunk = int(chunk_steps)
        self.clock = clock or time.monotonic
        self.queue = collections.deque()
        self.active = []                 # (request, member, footprint, admitted, steps_done, nt)
        self.used = 0
        self.batch = None
        self.inflight = 0                # steps of the chunk running on the device
        self.results = []
        self.chunks = 0
        self.peak_used = 0
        self.peak_slots = 0
        self.cancelled = set()           # running requests to withdraw at the next boundary
        self.retiring = []               # left the active set; results not yet collected
        self.to_free = []                # results collected; memory freed at the next boundary
        self.joined = False              # admitted since the batch was built
        self.running = []                # entries stepped by the chunk in flight
        # host seconds spent at chunk boundaries, by phase
        self.phase_s = dict.fromkeys(('wait', 'retire', 'admit', 'batch', 'prepare'), 0.0)

    # -- memory ---------------------------------------------------------------
    def _need(self, req):
        """(bytes the member will hold while active, bytes its admission adds
        to the device: less when a kept simulation or a resident beam is
        reused)."""
        new = self.backend.footprint(req)
        mb = getattr(self.backend, 'model_bytes', None)
        return (mb(req) if mb else new), new

    def cached_bytes(self):
        f = getattr(self.backend, 'cached_bytes', None)
        return f() if f else 0

    def in_use(self):
        """HBM held now: active members, members awaiting their free, and
        every resident beam and kept simulation."""
        return self.used + self.cached_bytes()

    # -- queue ----------------------------------------------------------------
    def submit(self, req: Request):
        check = getattr(self.backend, 'check', None)
        if check:
            check(req)                      # ValueError on malformed inputs
        fp = self.backend.footprint(req)
        if fp > self.budget:
            raise AdmissionError(f'request {req.rid} needs {fp} bytes, budget {self.budget}')
        self.queue.append(req)

    def cancel(self, rid):
        """Withdraw request ``rid``: a queued one at once, a running one at
        the next chunk boundary (its device work in flight completes first).
        Its result carries error 'cancelled'.  False if it is not here."""
        for req in self.queue:
            if req.rid == rid:
                self.queue.remove(req)
                now = self.clock()
                self.results.append(Result(rid, req.name, req.arrival, now, now, 0, None,
                                           'cancelled'))
                return True
        if any(e[0].rid == rid for e in self.active):
            self.cancelled.add(rid)
            return True
        return False

    # -- one boundary + chunk --------------------------------------------------
    def _retire(self):
        """Take finished and cancelled members out of the active set; their
        results are collected after the next chunk is launched and their
        memory is freed at the boundary after that (_collect)."""
        keep = []
        for entry in self.active:
            req = entry[0]
            if req.rid in self.cancelled or entry[4] >= entry[5]:
                self.retiring.append(entry)
            else:
                keep.append(entry)
        changed = len(keep) != len(self.active)
        self.active = keep
        return changed

    def _collect(self):
        """Far field and sigma0 of the retiring members (device work on their
        own streams, overlapping the chunk in flight), their results, and
        their release deferred to the next boundary."""
        for entry in self.retiring:
            req, member, fp, t_admit, done, nt = entry
            if req.rid in self.cancelled:
                self.cancelled.discard(req.rid)
                self.results.append(Result(req.rid, req.name, req.arrival, t_admit,
                                           self.clock(), done, None, 'cancelled'))
            else:
                out, err = None, None
                try:
                    out = self.backend.collect(member)
                except Exception as e:          # a failing request does not stop the loop
                    err = f'{type(e).__name__}: {e}'
                self.results.append(Result(req.rid, req.name, req.arrival, t_admit,
                                           self.clock(), nt, out, err))
            self.to_free.append(entry)
        self.retiring = []

    def _free(self):
        for req, member, fp, _, _, _ in self.to_free:
            self.backend.release(member)
            self.used -= fp
        self.to_free = []

    def _device_idle(self):
        """True when the chunk in flight has finished: host work should stop
        so that the next chunk starts (backends without a query: never)."""
        busy = getattr(self.backend, 'busy', None)
        return bool(self.inflight and busy is not None and not busy(self.batch))

    def _admit(self, first_only=False):
        """Admit queued requests in arrival order while slots and memory
        allow; stop early once the device has run out of work (the next
        chunk should start first), or after one admission when nothing runs
        yet (``first_only``)."""
        changed = False
        while self.queue and len(self.active) < self.max_slots:
            if changed and (first_only or self._device_idle()):
                break
            req = self.queue[0]
            fp, new = self._need(req)
            over = self.in_use() + new - self.budget
            if over > 0:
                evict = getattr(self.backend, 'evict_cached', None)
                if evict:
                    evict(over, req)            # kept simulations, unreferenced beams
                    fp, new = self._need(req)
                    over = self.in_use() + new - self.budget
            if over > 0:
                break                           # FIFO: wait for memory to free up
            self.queue.popleft()
            try:
                member = self.backend.admit(req)
            except Exception as e:
                now = self.clock()
                self.results.append(Result(req.rid, req.name, req.arrival, now, now, 0, None,
                                           f'{type(e).__name__}: {e}'))
                continue
            self.used += fp
            self.active.append([req, member, fp, self.clock(), 0, self.backend.steps(req)])
            changed = True
        self.peak_used = max(self.peak_used, self.in_use())
        self.peak_slots = max(self.peak_slots, len(self.active))
        return changed

    def _finish_chunk(self):
        if self.inflight:
            self.backend.sync(self.batch)
            for e in self.running:              # members admitted since did not step
                e[4] = min(e[5], e[4] + self.inflight)
            self.inflight = 0
            self.running = []

    def _prepare_queued(self):
        prepare = getattr(self.backend, 'prepare', None)
        if prepare is None:
            return
        free = max(1, self.max_slots - len(self.active))
        for req in list(self.queue)[:free]:
            if self._device_idle():
                break
            try:
                prepare(req)
            except Exception:               # admit() repeats the work and reports it
                pass

    def step(self):
        """One boundary: wait for the chunk in flight, free the members
        retired at the previous boundary, retire finished and cancelled
        members, start the next chunk of the remaining ones, and while it
        runs admit queued requests (they join at the next boundary), prepare
        the ones still queued and collect the retired members' results.
        Returns the number of active members."""
        ph = self.phase_s
        t0 = time.perf_counter()
        self._finish_chunk()
        t1 = time.perf_counter()
        self._free()
        changed = self._retire() or self.joined
        self.joined = False
        t2 = time.perf_counter()
        if not self.active:
            # nothing to run: admit one, start it, admit more while it runs
            changed = self._admit(first_only=True) or changed
        t3 = time.perf_counter()
        if changed and self.batch is not None:
            self.backend.drop_batch(self.batch)
            self.batch = None
        if self.active:
            if self.batch is None:
                self.batch = self.backend.make_batch([e[1] for e in self.active])
            n = min(self.chunk, max(e[5] - e[4] for e in self.active))
            self.backend.run(self.batch, n)        # asynchronous
            self.inflight = n
            self.running = list(self.active)
            self.chunks += 1
        t4 = time.perf_counter()
        self.joined = self._admit()
        t5 = time.perf_counter()
        self._prepare_queued()
        t6 = time.perf_counter()
        # last: reading a far field back may wait for the chunk's kernels
        self._collect()
        t7 = time.perf_counter()
        ph['wait'] += t1 - t0
        ph['retire'] += (t2 - t1) + (t7 - t6)
        ph['admit'] += (t3 - t2) + (t5 - t4)
        ph['batch'] += t4 - t3
        ph['prepare'] += t6 - t5
        return len(self.active)

    def idle(self):
        return (not self.queue and not self.active and not self.retiring
                and not self.to_free and not self.inflight)

    def drain(self):
        while not self.idle():
            self.step()
        self.close()

    def close(self):
        self._finish_chunk()
        if self.batch is not None:
            self.backend.drop_batch(self.batch)
            self.batch = None

    def serve(self, requests, time_scale=1.0):
        """Replay ``requests`` (sorted by arrival, seconds from start) against
        the server clock: each is submitted once its arrival time has passed;
        the loop runs until all are done.  Returns the results."""
        t0 = self.clock()
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        while pending or not self.idle():
            now = self.clock() - t0
            while pending and pending[0].arrival * time_scale <= now:
                r = pending.popleft()
                r.arrival = t0 + r.arrival * time_scale
                self.submit(r)
            if self.idle() and pending:
                # nothing to do until the next arrival
                wait = pending[0].arrival * time_scale - (self.clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                continue
            self.step()
        self.close()
        return self.results


def latency_stats(results):
    lat = np.array([r.latency for r in results if r.error is None])
    if lat.size == 0:
        return {}
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {'count': int(lat.size), 'p50_s': float(np.percentile(lat, 50)),
            'p95_s': float(np.percentile(lat, 95)), 'p99_s': float(np.percentile(lat, 99)),
            'mean_s': float(lat.mean()), 'span_s': float(span),
            'models_per_hour': float(lat.size / span * 3600.0) if span > 0 else None}
