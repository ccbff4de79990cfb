"""Batched and served runs on the device (needs an MI355X):

* every member of a continuously-batched mix gives bit-identical phasors
  and sigma0 to the same request run alone;
* end to end (incident beam, main run, far field, sigma0) the FP32 device
  pipeline matches the FP64 oracle's backscatter sigma0 within 0.05 dB, for
  the plain workload, a batched soil realization (HH and VV) and a served
  canopy request.
"""
import copy

import numpy as np
import pytest

import helpers as H
import oracle
from sar_pixel_simulator_amd import _native as native
from sar_pixel_simulator_amd import workloads
from sar_pixel_simulator_amd.hip import batch as B
from sar_pixel_simulator_amd.hip.serve import HipBackend, Request, Server, latency_stats
from sar_pixel_simulator_amd.materials import compute_material_arrays, synthetic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _request_params(k):
    sizes = [(33, 31, 35, 14), (45, 39, 37, 15), (37, 41, 33, 13), (41, 35, 39, 16)]
    pols = ['HH', 'VV', 'HH', 'VV']
    thetas = [20.0, 30.0, 40.0, 20.0]
    nx, ny, nz, nzm = sizes[k]
    pc = workloads.base_params(polarization=pols[k], theta=thetas[k], incident_duration=2.0,
                               illumination_width=1.0, simulation_duration=3.0)
    workloads.explicit_geometry(pc, nx, ny, nz, nzm, 0.05)
    g = pc.geom_params
    m = synthetic.rough_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                             workloads.lam_of(pc), nzm - 4, 100 + k)
    return pc, m


def test_serving_device_built_materials_mix_equals_alone():
    """Requests that carry MaterialSpec's (built on the device at
    admission): each result in the mix equals the same spec run alone."""
    from sar_pixel_simulator_amd.materials import device
    say('case: 3 spec requests alone, then served together')
    native.init(0)
    beams = B.BeamCache()
    specs = []
    for k in range(3):
        pc, _ = _request_params(k)
        g = pc.geom_params
        specs.append((pc, device.canopy_over_soil_spec(pc, g.size_z_material // 2, seed=50 + k,
                                                       density=20.0)))
    alone = {}
    for k, (pc, spec) in enumerate(specs):
        dm = B.DeviceModel.from_spec(pc, spec, beams, name=f's{k}')
        B.run_plain([dm])
        alone[k] = dm.sigma0()
        dm.release()
    be = HipBackend(0)
    be.beams = beams
    srv = Server(be, hbm_budget=64 << 30, max_slots=2, chunk_steps=16)
    for k, (pc, spec) in enumerate(specs):
        srv.submit(Request(rid=k, params=pc, spec=spec, name=f's{k}', arrival=0.0))
    srv.drain()
    assert len(srv.results) == 3
    for r in srv.results:
        assert r.error is None, r.error
        for key in ('sigma_h', 'sigma_v'):
            assert np.array_equal(r.output[key], alone[r.rid][key]), (r.rid, key)
    beams.release()


def test_serving_mix_equals_alone():
    say('case: 4 requests alone, then served together')
    native.init(0)
    beams = B.BeamCache()
    alone = {}
    for k in range(4):
        pc, m = _request_params(k)
        dm = B.DeviceModel.from_material(pc, m, beams, name=f'r{k}')
        B.run_plain([dm])
        alone[k] = (dm.sim.phasors(), dm.sigma0())
        dm.release()
    be = HipBackend(0)
    be.beams = beams
    srv = Server(be, hbm_budget=64 << 30, max_slots=3, chunk_steps=16)
    reqs = []
    for k in range(4):
        pc, m = _request_params(k)
        reqs.append(Request(rid=k, params=pc, material=m, name=f'r{k}', arrival=0.0))
    srv.submit(reqs[0])
    srv.step()
    srv.submit(reqs[1])        # joins while r0 runs
    srv.step()
    srv.submit(reqs[2])
    srv.submit(reqs[3])        # waits for a slot (max 3)
    srv.drain()
    assert len(srv.results) == 4
    for r in srv.results:
        assert r.error is None, r.error
        want_ph, want_s = alone[r.rid]
        got = r.output
        for key in ('sigma_h', 'sigma_v'):
            assert np.array_equal(got[key], want_s[key]), (r.rid, key)
    assert srv.peak_slots == 3
    st = latency_stats(srv.results)
    say('latency', st)
    beams.release()


def test_batched_matches_plain_phasors():
    say('case: 6 models batched vs plain')
    native.init(0)
    beams = B.BeamCache()
    cases = [_request_params(k % 4) for k in range(6)]
    plain = []
    for pc, m in cases:
        dm = B.DeviceModel.from_material(pc, m, beams)
        B.run_plain([dm])
        plain.append(dm.sim.phasors())
        dm.release()
    models = [B.DeviceModel.from_material(pc, m, beams) for pc, m in cases]
    B.run_batched(models, chunk=24)
    for k, dm in enumerate(models):
        ph = dm.sim.phasors()
        for n in native.PHASOR_NAMES:
            assert np.array_equal(ph[n], plain[k][n]), (k, n)
        dm.release()
    beams.release()


def _db(x):
    return 10 * np.log10(x)


def _fp64_sigma0(pc32, material):
    """The FP64 oracle pipeline of one model: incident run, main run, top-face
    far field and sigma (h, v) over theta = -80..80."""
    p64 = copy.deepcopy(pc32)
    p64.processing_params.datatype = 'float64'
    g = p64.geom_params
    pex, pey = H.excitation(p64)
    rec, inc = H.oracle_incident(p64, pex, pey)
    mm = copy.deepcopy(material)
    mm.relative_permittivity = np.asarray(mm.relative_permittivity, np.float64)
    mm.conductivity = np.asarray(mm.conductivity, np.float64)
    Ca, Cb = compute_material_arrays(p64, mm)
    sim = H.oracle_main(p64, Ca, Cb, rec)
    n = g.size_pml
    nxp, nyp, nzp = g.size_x - 2 * n, g.size_y - 2 * n, g.size_z - 2 * n
    th = np.linspace(-80.0, 80.0, 161)
    ff = oracle.far_field(sim.phasors(), nxp, nyp, nzp, g.cell_size, 2000.0, th, [0.0],
                          p64.inc_beam_params.frequency)
    return oracle.sigma0(ff, oracle.incident_power(inc, g.cell_size), nxp, nyp, g.cell_size)


def _backscatter(sigmas, pc):
    """sigma at theta = -theta_inc of the co-polarised channel."""
    sh, sv = sigmas
    i = 80 - int(round(pc.inc_beam_params.theta_angle))
    return (sh if pc.inc_beam_params.polarization == 'HH' else sv)[i]


def test_sigma0_vs_fp64_oracle():
    """Plain workload (81^3 cells, 480 steps, rough soil, HH, theta 20):
    device FP32 pipeline vs the FP64 oracle pipeline, backscatter sigma0."""
    say('case: sigma0 81^3 x 480 steps vs FP64 oracle')
    pc = workloads.plain_params()
    m = workloads.plain_material(pc, seed=0)
    native.init(0)
    beams = B.BeamCache()
    dm = B.DeviceModel.from_material(pc, m, beams)
    B.run_plain([dm])
    s = dm.sigma0()
    dm.release()
    beams.release()
    got_h = s['backscatter'][0]
    sh, sv = _fp64_sigma0(pc, m)
    want_h = sh[60]                                   # theta = -20
    say('sigma0 HH dB: device', _db(got_h), 'fp64 oracle', _db(want_h))
    assert abs(_db(got_h) - _db(want_h)) <= 0.05
    # the whole backscatter half-plane, not just one angle
    dev = _db(s['sigma_h'][20:80])
    ref = _db(sh[20:80])
    assert np.max(np.abs(dev - ref)) <= 0.05


def test_sigma0_batched_members_vs_fp64_oracle():
    """Batched workload: the smallest soil realization (3 x 3 lambda at 0.03
    lambda, 800 steps), HH and VV, stepped together as one batch; each
    member's backscatter sigma0 vs its FP64 oracle pipeline."""
    native.init(0)
    cases = workloads.batched_params_list(count=1)
    mats = [workloads.batched_material(pc, i) for i, pc in cases]
    g = cases[0][1].geom_params
    say('case: batched members', g.size_x, g.size_y, g.size_z, 'x', g.size_t, 'steps')
    beams = B.BeamCache()
    models = [B.DeviceModel.from_material(pc, m, beams) for (i, pc), m in zip(cases, mats)]
    B.run_batched(models, chunk=64)
    got = [m.sigma0() for m in models]
    for m in models:
        m.release()
    beams.release()
    for ((i, pc), m, s) in zip(cases, mats, got):
        want = _backscatter(_fp64_sigma0(pc, m), pc)
        dev = s['backscatter'][0 if pc.inc_beam_params.polarization == 'HH' else 1]
        say(pc.inc_beam_params.polarization, 'sigma0 dB: device', _db(dev), 'fp64 oracle', _db(want))
        assert abs(_db(dev) - _db(want)) <= 0.05


def test_sigma0_served_request_vs_fp64_oracle():
    """Serving workload: a canopy-over-soil request (4 x 4 lambda, 2 lambda
    high, 0.025 lambda) served in a mix with a second request; its
    backscatter sigma0 vs the FP64 oracle pipeline."""
    from sar_pixel_simulator_amd.hip.measure import serving_requests
    native.init(0)
    reqs = serving_requests(2, rate_hz=1.0, lateral=(4.0,), height=(2.0,))
    g = reqs[0].params.geom_params
    say('case: served request', g.size_x, g.size_y, g.size_z, 'x', g.size_t, 'steps')
    be = HipBackend(0)
    srv = Server(be, hbm_budget=64 << 30, max_slots=2, chunk_steps=64)
    for r in reqs:
        r.arrival = 0.0
        srv.submit(r)
    srv.drain()
    res = {r.rid: r for r in srv.results}
    assert res[0].error is None, res[0].error
    be.beams.release()
    pc = reqs[0].params
    want = _backscatter(_fp64_sigma0(pc, reqs[0].material), pc)
    out = res[0].output
    dev = out['backscatter'][0 if pc.inc_beam_params.polarization == 'HH' else 1]
    say(pc.inc_beam_params.polarization, 'theta', pc.inc_beam_params.theta_angle,
        'sigma0 dB: device', _db(dev), 'fp64 oracle', _db(want))
    assert abs(_db(dev) - _db(want)) <= 0.05


def test_recycled_simulation_equals_fresh():
    """A simulation kept in the pool after one model and reused for the next
    model on the same grid (another material, another polarisation and
    angle, so another beam) gives the fresh simulation's phasors and sigma0
    bit for bit; a host-volume model after a device-built one too."""
    from sar_pixel_simulator_amd.materials import device
    say('case: recycled simulations vs fresh')
    native.init(0)
    beams = B.BeamCache()
    pool = B.SimPool()
    pc0, m0 = _request_params(0)
    pc1 = copy.deepcopy(pc0)
    pc1.inc_beam_params.polarization = 'VV'
    pc1.inc_beam_params.theta_angle = 40.0
    g = pc1.geom_params
    m1 = synthetic.rough_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                              workloads.lam_of(pc1), g.size_z_material - 3, 777)
    spec = device.canopy_over_soil_spec(pc0, g.size_z_material // 2, seed=31, density=20.0)
    fresh = []
    for pc, kind in ((pc0, 'spec'), (pc1, 'host')):
        dm = (B.DeviceModel.from_spec(pc, spec, beams) if kind == 'spec'
              else B.DeviceModel.from_material(pc, m1, beams))
        B.run_plain([dm])
        fresh.append((dm.sim.phasors(), dm.sigma0()))
        dm.release()
    first = B.DeviceModel.from_material(pc0, m0, beams, pool=pool)
    B.run_plain([first])
    first.release()                                  # kept
    assert pool.has(pc0) and pool.bytes() > 0
    for (want_ph, want_s), (pc, kind) in zip(fresh, ((pc0, 'spec'), (pc1, 'host'))):
        dm = (B.DeviceModel.from_spec(pc, spec, beams, pool=pool) if kind == 'spec'
              else B.DeviceModel.from_material(pc, m1, beams, pool=pool))
        assert not pool.has(pc)                      # the kept one was taken
        B.run_plain([dm])
        ph = dm.sim.phasors()
        for n in native.PHASOR_NAMES:
            assert np.array_equal(ph[n], want_ph[n]), (kind, n)
        got_s = dm.sigma0()
        for key in ('sigma_h', 'sigma_v'):
            assert np.array_equal(got_s[key], want_s[key]), (kind, key)
        dm.release()
    pool.release()
    beams.release()


def test_tuned_batch_equals_plain():
    """The batch autotuner steps the members in each candidate execution,
    resets them and runs the fastest: results equal each model run alone."""
    say('case: tuned batch of 4 vs plain')
    native.init(0)
    beams = B.BeamCache()
    cases = [_request_params(k) for k in range(4)]
    plain = []
    for pc, m in cases:
        dm = B.DeviceModel.from_material(pc, m, beams)
        B.run_plain([dm])
        plain.append(dm.sim.phasors())
        dm.release()
    models = [B.DeviceModel.from_material(pc, m, beams) for pc, m in cases]
    B.run_batched(models, chunk=24, mode='tune')
    for k, dm in enumerate(models):
        assert dm.sim.step_count() == dm.nt
        ph = dm.sim.phasors()
        for n in native.PHASOR_NAMES:
            assert np.array_equal(ph[n], plain[k][n]), (k, n)
        dm.release()
    beams.release()
