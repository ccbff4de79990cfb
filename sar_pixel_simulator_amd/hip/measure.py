"""Throughput and latency measurements of the batched and served paths.

* ``batched_vs_plain``: material-models/hour of the same set of models run
  one at a time (the reference's experiment loop) and as one device batch.
  Both count, per model, the Ca/Cb upload, all time steps and the far field
  plus sigma0 from the device phasor sums; the incident beams are generated
  beforehand (cached per beam key) and reported separately.
* ``serving``: a Poisson stream of requests through the continuous-batching
  server with an HBM budget; p50/p95/p99 latency and models/hour, against
  (a) the same requests run one at a time on the device and (b) a sample of
  them through the reference-style loop (host material volumes, Ca/Cb,
  MainSimHip.run writing the results file, process_results).
"""
import time

import numpy as np

from .. import _native as native
from .. import workloads
from ..materials import compute_material_arrays, synthetic
from . import batch as B
from . import common
from .serve import HipBackend, Request, Server, latency_stats


def _models_for(cases, beams):
    return [B.DeviceModel(pc, Ca, Cb, beams, name) for pc, Ca, Cb, name in cases]


def batched_workload(count, seed=1000):
    """Soil realizations of uneven lateral size {3,4,6,8} lambda, depth
    U[1,2] lambda, cell 0.03 lambda, HH and VV (``count`` seeds x 2 pols)."""
    cases = []
    for i, pc in workloads.batched_params_list(count=count, seed=seed):
        m = workloads.batched_material(pc, i)
        Ca, Cb = compute_material_arrays(pc, m)
        cases.append((pc, Ca, Cb, f'soil_{i}_{pc.inc_beam_params.polarization}'))
    return cases


def _batched_reference_loop(n, beams, seed=1000, log=print):
    """Models/hour of the reference-style loop (experiment.py:69-97 with the
    HIP backend) over the first ``n`` models of the batched workload: per
    model its host soil volume (workloads.batched_material, as the
    reference's soil generators yield models), compute_material_arrays,
    MainSimHip.run (upload, all steps, phasors to the host, results file)
    and process_results (far field, powers, sigma0 into the file).  The
    incident beams are the resident ones of ``beams`` (generated beforehand,
    as in every leg)."""
    import tempfile
    from ..incidentbeam import forget_resident, make_resident
    from ..processresults import process_results
    from ..parameters import results_filename
    from .main_sim import MainSimHip
    t_total = 0.0
    done = 0
    with tempfile.TemporaryDirectory(prefix='fdtd_refloop_') as proj:
        for i, pc in workloads.batched_params_list(count=(n + 1) // 2, seed=seed):
            if done == n:
                break
            pc.processing_params.project_dir = proj
            beam, ph = beams.lookup(pc)
            make_resident(pc, beam, ph)
            name = f'soil_{i}_{pc.inc_beam_params.polarization}'
            t0 = time.perf_counter()
            m = workloads.batched_material(pc, i)
            Ca, Cb = compute_material_arrays(pc, m)
            sim = MainSimHip(pc)
            sim.run(Ca, Cb, name, {})
            sim.close()
            process_results(results_filename(pc, name))
            t_total += time.perf_counter() - t0
            forget_resident(pc)
            done += 1
    out = {'models': done, 's_per_model': round(t_total / done, 4),
           'models_per_hour': round(done / t_total * 3600.0, 1)}
    log(f'batched_vs_plain: reference-style loop {out}')
    return out


def _batch_alg_bytes(cases):
    from .roofline import algorithmic_bytes
    alg = 0
    for pc, _, _, _ in cases:
        g = pc.geom_params
        a = algorithmic_bytes(g.size_x, g.size_y, g.size_z, g.size_pml, None)
        alg += a['h'] + a['e'] + a['dft']
    return alg


def batched_vs_plain(count=32, chunk=96, device=0, log=print, step_probe=True, reference_loop=16):
    """The batched workload (``count`` soil realizations x HH/VV) one model
    at a time and as one device batch: models/hour of each, identical
    results, and the batch's step rate in each execution mode (the bound
    on batched / plain that the kernels allow).  With ``reference_loop`` > 0
    that many of the models also run through the reference-style loop
    (host volume, Ca/Cb, MainSimHip.run writing the results file,
    process_results), against which the batch is compared with its host
    volumes counted."""
    native.init(device)
    t0 = time.perf_counter()
    cases = batched_workload(count)
    t_host = time.perf_counter() - t0
    beams = B.BeamCache()
    t0 = time.perf_counter()
    for pc, _, _, _ in cases:
        beams.warm(pc)
    t_beams = time.perf_counter() - t0
    cells = [pc.geom_params.size_x * pc.geom_params.size_y * pc.geom_params.size_z
             for pc, _, _, _ in cases]
    steps = [pc.geom_params.size_t for pc, _, _, _ in cases]
    log(f'batched_vs_plain: {len(cases)} models ({t_host:.1f} s host volumes), '
        f'{len(beams)} beams in {t_beams:.1f} s')

    # plain: one model at a time (upload, all steps, far field + sigma0, free)
    t0 = time.perf_counter()
    plain_back = []
    for pc, Ca, Cb, name in cases:
        m = B.DeviceModel(pc, Ca, Cb, beams, name)
        B.run_plain([m])
        plain_back.append(m.sigma0()['sigma_h'])
        m.release()
    t_plain = time.perf_counter() - t0
    log(f'batched_vs_plain: plain {t_plain:.2f} s')

    # batched: all resident, stepped together in the execution the tuner
    # picks for this batch (its measurement, a one-time cost per batch
    # composition kept in the tuning cache, is reported apart)
    from . import tuning
    t0 = time.perf_counter()
    models = _models_for(cases, beams)
    t1 = time.perf_counter()
    (mode, launch), tuned = tuning.autotune_batch(models, save=False)
    t_tune = time.perf_counter() - t1
    B.run_batched(models, chunk=chunk, launch=launch, mode=mode)
    batch_back = [m.sigma0()['sigma_h'] for m in models]
    t_batch = time.perf_counter() - t0 - t_tune
    label = min(tuned, key=tuned.get)
    log(f'batched_vs_plain: tuner picked {label} in {t_tune:.2f} s: '
        f'{ {k: round(v, 3) for k, v in tuned.items()} }')
    same = all(np.array_equal(a, b) for a, b in zip(plain_back, batch_back))
    log(f'batched_vs_plain: batched {t_batch:.2f} s, identical results: {same}')
    alg = _batch_alg_bytes(cases)
    modes = {}
    if step_probe:
        # steady-state step of the whole batch in each execution mode, and of
        # the members one at a time (their own launches, synchronised each)
        for m in models:
            m.sim.reset()
        for label, mode_k in (('packed', 'packed'), ('serial_1_lane', (native.BATCH_SERIAL, 0, 1)),
                              ('serial_4_lanes', (native.BATCH_SERIAL, 0, 4))):
            db = B.DeviceBatch(models, mode=mode_k)
            db.run(4)
            db.sync()
            t1 = time.perf_counter()
            db.run(32)
            db.sync()
            ms = (time.perf_counter() - t1) * 1e3 / 32
            db.destroy()
            modes[label] = {'ms_per_step': round(ms, 4), 'alg_TBs': round(alg / ms / 1e9, 3),
                            'alg_frac_of_8TBs': round(alg / ms / 1e9 / 8.0, 3)}
        t1 = time.perf_counter()
        for m in models:
            m.sim.run(32)
            m.sim.sync()
        ms = (time.perf_counter() - t1) * 1e3 / 32
        modes['one_at_a_time'] = {'ms_per_step': round(ms, 4), 'alg_TBs': round(alg / ms / 1e9, 3),
                                  'alg_frac_of_8TBs': round(alg / ms / 1e9 / 8.0, 3)}
        log(f'batched_vs_plain: step rates {modes}')
    for m in models:
        m.release()
    ref = _batched_reference_loop(min(reference_loop, len(cases)), beams, log=log) if reference_loop else None
    beams.release()
    n = len(cases)
    cu = float(np.dot(cells, steps))
    out = {'workload': f'{count} soil realizations x HH/VV = {n} simulations, lateral '
                       '{3,4,6,8} lambda, depth U[1,2] lambda, cell 0.03 lambda, 800 steps',
           'models': n, 'beams': len(set(map(B.beam_key, [c[0] for c in cases]))),
           'beam_setup_s': round(t_beams, 2), 'host_volume_setup_s': round(t_host, 2),
           'plain_s': round(t_plain, 3), 'batched_s': round(t_batch, 3),
           'plain_models_per_hour': round(n / t_plain * 3600, 1),
           'batched_models_per_hour': round(n / t_batch * 3600, 1),
           'speedup': round(t_plain / t_batch, 3),
           'batch_execution': label, 'batch_tune_s': round(t_tune, 2),
           'batch_tune_ms_per_step': {k: round(v, 4) for k, v in tuned.items()},
           'batched_Mcell_updates_per_s': round(cu / t_batch / 1e6, 1),
           'plain_Mcell_updates_per_s': round(cu / t_plain / 1e6, 1),
           'step_alg_bytes': int(alg),
           'identical_results': bool(same),
           'grid_cells_min_max': [int(min(cells)), int(max(cells))],
           'steps': int(max(steps))}
    if ref:
        # the reference-style loop generates each model's host volume inside
        # its loop; the batch's volumes were generated before it, so they are
        # counted here
        with_host = n / (t_batch + t_host) * 3600
        out['reference_loop'] = ref
        out['batched_models_per_hour_with_host_volumes'] = round(with_host, 1)
        out['batched_vs_reference_loop'] = round(with_host / ref['models_per_hour'], 3)
    if modes:
        out['step_rate_by_mode'] = modes
        best = min(v['ms_per_step'] for k, v in modes.items() if k != 'one_at_a_time')
        out['step_bound_batched_vs_plain'] = round(modes['one_at_a_time']['ms_per_step'] / best, 3)
    return out


#: Canopy request sizes: lateral U[4, 12] lambda and height U[2, 4] lambda,
#: drawn from discrete steps so that requests share incident beams (a beam
#: depends on the lateral grid, the angle and the polarisation).
CANOPY_LATERAL = (4.0, 6.0, 8.0, 10.0, 12.0)
CANOPY_HEIGHT = (2.0, 3.0, 4.0)


def serving_requests(n, rate_hz, seed=7, lateral=CANOPY_LATERAL, height=CANOPY_HEIGHT,
                     thetas=(20.0, 30.0, 40.0), pols=('HH', 'VV'), cell=0.025,
                     device_materials=False):
    """Poisson arrivals of vegetation-canopy-over-soil requests.  Lateral
    sizes and heights are drawn from a few discrete values so that beams are
    shared between requests (a beam is keyed by grid size, angle and
    polarisation).  With ``device_materials`` a request carries a
    MaterialSpec (surface map and cylinder list) that the device expands,
    instead of the (nz, ny, nx) permittivity and conductivity volumes."""
    from ..materials import device
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for k in range(n):
        t += rng.exponential(1.0 / rate_hz)
        lat = float(rng.choice(lateral))
        h = float(rng.choice(height))
        pc = workloads.base_params(polarization=str(rng.choice(pols)),
                                   theta=float(rng.choice(thetas)), incident_duration=10.0,
                                   illumination_width=2.0, simulation_duration=12.0,
                                   material_fraction=0.6)
        workloads.set_geom_params(pc, lat, lat, h, 12.0, cell, 'wavelength')
        g = pc.geom_params
        soil = max(2, g.size_z_material // 3)
        density = float(rng.uniform(2.0, 6.0))
        if device_materials:
            spec = device.canopy_over_soil_spec(pc, soil, seed=1000 + k, density=density)
            reqs.append(Request(rid=k, params=pc, spec=spec, name=f'req{k}', arrival=t))
        else:
            m = synthetic.canopy_over_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                                           workloads.lam_of(pc), soil, seed=1000 + k,
                                           density=density)
            reqs.append(Request(rid=k, params=pc, material=m, name=f'req{k}', arrival=t))
    return reqs


def material_setup(n=4, lateral=8.0, height=4.0, cell=0.025, device=0, log=print):
    """Time to have a canopy-over-soil model's Ca, Cb resident, per model:
    host path (generate the volumes, compute_material_arrays, upload) vs
    device path (surface map and cylinder list to the device builder), and
    the bytes each ships."""
    from ..materials import compute_material_arrays, device as dev
    native.init(device)
    pc = workloads.base_params(polarization='VV', theta=30.0, material_fraction=0.6)
    workloads.set_geom_params(pc, lateral, lateral, height, 12.0, cell, 'wavelength')
    g = pc.geom_params
    soil = max(2, g.size_z_material // 3)
    sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
    t_host = t_dev = 0.0
    host_bytes = dev_bytes = 0
    for k in range(n):
        t0 = time.perf_counter()
        m = synthetic.canopy_over_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                                       workloads.lam_of(pc), soil, seed=2000 + k, density=4.0)
        Ca, Cb = compute_material_arrays(pc, m)
        sim.set_materials(Ca, Cb)
        sim.sync()
        t_host += time.perf_counter() - t0
        host_bytes = Ca.nbytes + Cb.nbytes
        t0 = time.perf_counter()
        spec = dev.canopy_over_soil_spec(pc, soil, seed=2000 + k, density=4.0)
        sim.build_materials(*dev.to_c(spec, pc))
        sim.sync()
        t_dev += time.perf_counter() - t0
        dev_bytes = spec.payload_bytes()
    sim.destroy()
    out = {'grid': [g.size_x, g.size_y, g.size_z], 'models': n,
           'host_path_s_per_model': round(t_host / n, 4),
           'device_path_s_per_model': round(t_dev / n, 4),
           'speedup': round(t_host / t_dev, 2),
           'host_bytes_shipped': int(host_bytes), 'device_bytes_shipped': int(dev_bytes)}
    log(f'material_setup: {out}')
    return out


def _reference_loop(reqs, be, log):
    """Models/hour of the reference-style loop (experiment.py:69-97 with the
    HIP backend) over ``reqs``: per model the host material volumes
    (synthetic.canopy_over_soil), compute_material_arrays, MainSimHip.run
    (upload, all steps, phasors to the host, results file) and
    process_results (far field, powers, sigma0 into the file).  Incident
    beams are the ones already resident (generated beforehand, as in every
    leg)."""
    import tempfile
    from ..incidentbeam import forget_resident, make_resident
    from ..processresults import process_results
    from ..parameters import results_filename
    from .main_sim import MainSimHip
    t_total = 0.0
    with tempfile.TemporaryDirectory(prefix='fdtd_refloop_') as proj:
        for r in reqs:
            pc = r.params
            pc.processing_params.project_dir = proj
            beam, ph = be.beams.lookup(pc)
            make_resident(pc, beam, ph)
            g = pc.geom_params
            soil = max(2, g.size_z_material // 3)
            t0 = time.perf_counter()
            m = synthetic.canopy_over_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                                           workloads.lam_of(pc), soil, seed=5000 + r.rid,
                                           density=4.0)
            Ca, Cb = compute_material_arrays(pc, m)
            sim = MainSimHip(pc)
            sim.run(Ca, Cb, r.name, {})
            sim.close()
            process_results(results_filename(pc, r.name))
            t_total += time.perf_counter() - t0
            forget_resident(pc)
    out = {'models': len(reqs), 's_per_model': round(t_total / len(reqs), 4),
           'models_per_hour': round(len(reqs) / t_total * 3600.0, 1)}
    log(f'serving: reference-style loop {out}')
    return out


def serving(n=200, rate_factor=8.0, budget_gb=None, max_slots=256, chunk=96, device=0, log=print,
            device_materials=False, reference_loop=0, mode='auto'):
    """Serve ``n`` Poisson canopy requests through the continuous-batching
    server.  The same requests are first run one at a time (admit, all
    steps, sigma0, release): that plain service rate is the baseline (a),
    and the arrival rate is ``rate_factor`` times it, so that the queue
    grows until admission binds on HBM.  The budget defaults to the device
    memory free after the incident beams are resident (the beams are
    counted in it), less 2 GiB.  ``reference_loop`` > 0 times that many of
    the requests through the reference-style loop: baseline (b)."""
    native.init(device)
    t0 = time.perf_counter()
    reqs = serving_requests(n, rate_hz=1.0, device_materials=device_materials)
    t_inputs = time.perf_counter() - t0
    be = HipBackend(device, mode=mode)
    t0 = time.perf_counter()
    for r in reqs:                         # beams first: reported separately
        be.beams.warm(r.params)
    t_beams = time.perf_counter() - t0
    log(f'serving: {n} requests ({t_inputs:.1f} s inputs), {len(be.beams)} beams in {t_beams:.1f} s')
    ref = _reference_loop(reqs[:reference_loop], be, log) if reference_loop else None
    # baseline (a): the same requests one at a time
    t0 = time.perf_counter()
    for k, r in enumerate(reqs):
        m = be.admit(r)
        B.run_plain([m])
        m.sigma0()
        m.release()
        if k % 50 == 49:
            log(f'serving: plain {k + 1}/{n} in {time.perf_counter() - t0:.1f} s')
    t_plain = time.perf_counter() - t0
    svc = t_plain / n
    rate_hz = rate_factor / svc
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    t = 0.0
    for r, gap in zip(reqs, gaps):       # rescale the Poisson stream to the rate
        t += gap / rate_hz
        r.arrival = t
    if budget_gb:
        budget = int(budget_gb * (1 << 30))
    else:
        budget = int(native.device_info()['free_mem'] + be.beams.resident_bytes() - (2 << 30))
    srv = Server(be, hbm_budget=budget, max_slots=max_slots, chunk_steps=chunk)
    t0 = time.perf_counter()
    res = srv.serve(reqs)
    wall = time.perf_counter() - t0
    be.beams.release()
    st = latency_stats(res)
    cells = [r.params.geom_params.size_x * r.params.geom_params.size_y * r.params.geom_params.size_z
             for r in reqs]
    st.update({'workload': f'{n} Poisson canopy-over-soil requests, lateral '
                           f'{CANOPY_LATERAL[0]:g}-{CANOPY_LATERAL[-1]:g} lambda, height '
                           f'{CANOPY_HEIGHT[0]:g}-{CANOPY_HEIGHT[-1]:g} lambda, cell 0.025 lambda, '
                           'theta {20,30,40}, HH/VV, 960 steps',
               'requests': n, 'errors': sum(r.error is not None for r in res),
               'materials': 'built on the device' if device_materials else 'host volumes uploaded',
               'grid_cells_min_max': [int(min(cells)), int(max(cells))],
               'beams': len(set(B.beam_key(r.params) for r in reqs)), 'beam_setup_s': round(t_beams, 2),
               'arrival_rate_hz': round(rate_hz, 3), 'rate_factor_vs_plain': rate_factor,
               'hbm_budget_bytes': budget, 'peak_hbm_bytes': srv.peak_used,
               'peak_hbm_frac_of_budget': round(srv.peak_used / budget, 3),
               'peak_members': srv.peak_slots, 'chunks': srv.chunks, 'wall_s': round(wall, 2),
               'boundary_host_s': {k: round(v, 3) for k, v in srv.phase_s.items()},
               'plain_s': round(t_plain, 2),
               'plain_models_per_hour': round(3600.0 / svc, 1)})
    if st.get('models_per_hour'):
        st['served_vs_plain'] = round(st['models_per_hour'] * svc / 3600.0, 3)
    if ref:
        st['reference_loop'] = ref
        if st.get('models_per_hour'):
            st['served_vs_reference_loop'] = round(st['models_per_hour'] / ref['models_per_hour'], 3)
    log(f'serving: {st}')
    return st


def batch_step_rate(count=32, steps=30, mode='auto', device=0):
    """Steady-state step rate of the batched workload as one batch: all
    members resident, ``steps`` steps after a short warm-up, in batch mode
    ``mode`` ('auto', 'packed' or a fdtd_batch_set_mode tuple).  Returns ms
    per step, Gcell-updates/s and the algorithmic rate (the split kernels'
    and the DFT's compulsory bytes per step over the step)."""
    native.init(device)
    cases = batched_workload(count)
    beams = B.BeamCache()
    models = _models_for(cases, beams)
    alg = _batch_alg_bytes(cases)
    cells = sum(pc.geom_params.size_x * pc.geom_params.size_y * pc.geom_params.size_z
                for pc, _, _, _ in cases)
    db = B.DeviceBatch(models, mode=mode)
    db.run(5)
    db.sync()
    t0 = time.perf_counter()
    db.run(steps)
    db.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    used = db.mode
    db.destroy()
    for m in models:
        m.release()
    beams.release()
    return {'models': len(cases), 'cells': cells, 'steps': steps, 'mode': list(used),
            'ms_per_step': round(ms, 4), 'Gcell_updates_per_s': round(cells / ms / 1e6, 1),
            'alg_TBs': round(alg / ms / 1e9, 3)}


if __name__ == '__main__':
    import json
    import sys
    what = sys.argv[1:] or ['batched', 'serving']
    out = {}
    if 'batched' in what:
        out['batched'] = batched_vs_plain()
    if 'serving' in what:
        out['serving'] = serving()
    if 'batch_step' in what:
        # batch_step [count] [packed]
        rest = what[what.index('batch_step') + 1:]
        count = int(rest[0]) if rest and rest[0].isdigit() else 32
        mode = 'packed' if 'packed' in rest else 'auto'
        out['batch_step'] = batch_step_rate(count=count, mode=mode)
    print(json.dumps(out))
