"""Scheduling logic of the serving loop against a fake device backend: HBM
admission, slots, FIFO order, continuous joining at chunk boundaries, error
isolation and latency statistics.  (Device results of the same loop are
checked on the GPU in test_gpu_batch_serve.py.)"""
import types

import pytest

from sar_pixel_simulator_amd.hip.serve import (AdmissionError, Request, Server,
                                               latency_stats)


class FakeMember:
    def __init__(self, req, nt):
        self.req, self.nt, self.steps = req, nt, 0


class FakeBackend:
    def __init__(self, sizes, steps, fail_collect=(), fail_admit=()):
        self.sizes, self.nsteps = sizes, steps
        self.fail_collect, self.fail_admit = set(fail_collect), set(fail_admit)
        self.batches = []
        self.released = []

    def footprint(self, req):
        return self.sizes[req.rid]

    def steps(self, req):
        return self.nsteps[req.rid]

    def admit(self, req):
        if req.rid in self.fail_admit:
            raise MemoryError('device allocation failed')
        return FakeMember(req, self.nsteps[req.rid])

    def make_batch(self, members):
        b = types.SimpleNamespace(members=list(members), runs=[])
        self.batches.append(b)
        return b

    def run(self, batch, n):
        batch.runs.append(n)
        for m in batch.members:
            m.steps = min(m.nt, m.steps + n)

    def sync(self, batch):
        pass

    def drop_batch(self, batch):
        pass

    def collect(self, m):
        if m.req.rid in self.fail_collect:
            raise RuntimeError('far field failed')
        return {'rid': m.req.rid, 'steps_run': m.steps}

    def release(self, m):
        self.released.append(m.req.rid)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1.0
        return self.t


def _reqs(n):
    return [Request(rid=i, params=None, name=f'r{i}') for i in range(n)]


def test_all_complete_with_exact_steps():
    n = 7
    be = FakeBackend({i: 10 for i in range(n)}, {i: 30 + 7 * i for i in range(n)})
    s = Server(be, hbm_budget=1000, max_slots=64, chunk_steps=16, clock=Clock())
    for r in _reqs(n):
        s.submit(r)
    s.drain()
    assert sorted(r.rid for r in s.results) == list(range(n))
    for r in s.results:
        assert r.error is None and r.output['steps_run'] == 30 + 7 * r.rid
    assert sorted(be.released) == list(range(n))
    assert s.used == 0


def test_budget_and_slots_bind():
    n = 9
    be = FakeBackend({i: 40 for i in range(n)}, {i: 20 for i in range(n)})
    s = Server(be, hbm_budget=100, max_slots=64, chunk_steps=8, clock=Clock())
    for r in _reqs(n):
        s.submit(r)
    s.drain()
    assert s.peak_used <= 100 and s.peak_slots == 2
    assert all(len(b.members) <= 2 for b in be.batches)
    be2 = FakeBackend({i: 1 for i in range(n)}, {i: 20 for i in range(n)})
    s2 = Server(be2, hbm_budget=100, max_slots=3, chunk_steps=8, clock=Clock())
    for r in _reqs(n):
        s2.submit(r)
    s2.drain()
    assert s2.peak_slots == 3 and len(s2.results) == n


def test_fifo_and_oversize_rejected():
    be = FakeBackend({0: 60, 1: 60, 2: 10, 3: 500}, {0: 8, 1: 8, 2: 8, 3: 8})
    s = Server(be, hbm_budget=100, chunk_steps=8, clock=Clock())
    with pytest.raises(AdmissionError):
        s.submit(Request(rid=3, params=None))
    for r in _reqs(3):
        s.submit(r)
    s.drain()
    admitted = sorted(s.results, key=lambda r: r.admitted)
    # request 2 fits next to 0 but waits behind 1 (arrival order is kept)
    assert [r.rid for r in admitted] == [0, 1, 2]


def test_late_arrival_joins_running_batch():
    be = FakeBackend({0: 10, 1: 10}, {0: 64, 1: 16})
    s = Server(be, hbm_budget=100, chunk_steps=8, clock=Clock())
    s.submit(Request(rid=0, params=None))
    s.step()
    s.step()
    s.submit(Request(rid=1, params=None))   # arrives while 0 is running
    s.step()                                # admitted while 0's next chunk runs
    assert [m.req.rid for m in be.batches[-1].members] == [0]
    assert [e[0].rid for e in s.active] == [0, 1]
    s.step()                                # joins at the following boundary
    assert [m.req.rid for m in be.batches[-1].members] == [0, 1]
    s.drain()
    res = {r.rid: r for r in s.results}
    assert res[1].finished < res[0].finished      # the short late one overtakes
    assert res[0].output['steps_run'] == 64 and res[1].output['steps_run'] == 16


def test_failures_are_isolated():
    be = FakeBackend({i: 1 for i in range(4)}, {i: 8 for i in range(4)},
                     fail_collect={1}, fail_admit={2})
    s = Server(be, hbm_budget=100, chunk_steps=8, clock=Clock())
    for r in _reqs(4):
        s.submit(r)
    s.drain()
    res = {r.rid: r for r in s.results}
    assert res[1].error.startswith('RuntimeError') and res[2].error.startswith('MemoryError')
    assert res[0].error is None and res[3].error is None
    assert 1 in be.released and 2 not in be.released


def test_latency_stats():
    be = FakeBackend({i: 1 for i in range(5)}, {i: 8 * (i + 1) for i in range(5)})
    s = Server(be, hbm_budget=100, chunk_steps=8, clock=Clock())
    for r in _reqs(5):
        s.submit(r)
    s.drain()
    st = latency_stats(s.results)
    assert st['count'] == 5
    assert st['p50_s'] <= st['p95_s'] <= st['p99_s']
    assert st['models_per_hour'] > 0


class LoggingBackend(FakeBackend):
    """FakeBackend that logs device launches, syncs and host preparation."""

    def __init__(self, *a, fail_prepare=(), **kw):
        super().__init__(*a, **kw)
        self.log, self.fail_prepare = [], set(fail_prepare)

    def run(self, batch, n):
        self.log.append(('run', tuple(m.req.rid for m in batch.members)))
        super().run(batch, n)

    def sync(self, batch):
        self.log.append(('sync',))

    def prepare(self, req):
        self.log.append(('prepare', req.rid))
        if req.rid in self.fail_prepare:
            raise ValueError('bad material')
        req.prepared = True


def test_queued_requests_are_prepared_while_a_chunk_runs():
    # one slot: request 1 waits while 0 runs, and is prepared between the
    # launch of 0's chunk and its sync
    be = LoggingBackend({0: 10, 1: 10}, {0: 16, 1: 8})
    s = Server(be, hbm_budget=100, max_slots=1, chunk_steps=8, clock=Clock())
    for r in _reqs(2):
        s.submit(r)
    s.drain()
    log = be.log
    first_run = log.index(('run', (0,)))
    assert log[first_run + 1] == ('prepare', 1) and log[first_run + 2] == ('sync',)
    assert log.index(('prepare', 1)) < log.index(('run', (1,)))
    assert [r.rid for r in sorted(s.results, key=lambda r: r.finished)] == [0, 1]
    assert all(r.error is None for r in s.results)
    assert log[-1] == ('sync',)                      # nothing left in flight
    assert s.inflight == 0


def test_failed_preparation_is_left_to_admission():
    be = LoggingBackend({0: 10, 1: 10, 2: 10}, {0: 8, 1: 8, 2: 8}, fail_prepare={1},
                        fail_admit={1})
    s = Server(be, hbm_budget=100, max_slots=1, chunk_steps=8, clock=Clock())
    for r in _reqs(3):
        s.submit(r)
    s.drain()
    res = {r.rid: r for r in s.results}
    assert res[1].error.startswith('MemoryError')     # reported by admit, not prepare
    assert res[0].error is None and res[2].error is None


def test_hip_backend_prepare_computes_host_coefficients():
    # prepare() needs no device: it is called unbound here
    import numpy as np
    from sar_pixel_simulator_amd import workloads
    from sar_pixel_simulator_amd.hip.serve import HipBackend
    from sar_pixel_simulator_amd.materials import compute_material_arrays, synthetic
    pc = workloads.small_params(nx=25, ny=23, nz=27, nz_mat=9)
    g = pc.geom_params
    m = synthetic.rough_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                             workloads.lam_of(pc), 5, 3)
    req = Request(rid=0, params=pc, material=m)
    HipBackend.prepare(None, req)
    Ca, Cb = compute_material_arrays(pc, m)
    assert req.prepared and np.array_equal(req.Ca, Ca) and np.array_equal(req.Cb, Cb)
    # arrays the caller gave, and device specs, are left alone
    given = Request(rid=1, params=pc, Ca=Ca, Cb=Cb)
    HipBackend.prepare(None, given)
    assert not given.prepared and given.Ca is Ca
    spec = Request(rid=2, params=pc, spec=object())
    HipBackend.prepare(None, spec)
    assert not spec.prepared and spec.Ca is None


class BusyBackend(FakeBackend):
    """FakeBackend whose device reports the chunk as finished (busy False)
    or still running (busy True)."""

    def __init__(self, *a, busy=True, **kw):
        super().__init__(*a, **kw)
        self.is_busy = busy
        self.queries = 0

    def busy(self, batch):
        self.queries += 1
        return self.is_busy


@pytest.mark.parametrize('busy', [True, False])
def test_host_work_only_while_the_device_is_busy(busy):
    # nothing runs yet: one request is admitted and started first; the rest
    # are admitted while its chunk runs -- all of them while the device
    # reports work, one per boundary once it reports none
    be = BusyBackend({i: 1 for i in range(5)}, {i: 64 for i in range(5)}, busy=busy)
    s = Server(be, hbm_budget=100, chunk_steps=8, clock=Clock())
    for r in _reqs(5):
        s.submit(r)
    s.step()
    assert [m.req.rid for m in be.batches[0].members] == [0]
    assert len(s.active) == (5 if busy else 2)
    s.drain()
    assert sorted(r.rid for r in s.results) == list(range(5))
    assert all(r.error is None and r.output['steps_run'] == 64 for r in s.results)
