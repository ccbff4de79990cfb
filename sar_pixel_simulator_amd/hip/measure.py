"""Throughput and latency measurements of the batched and served paths.

* ``batched_vs_plain``: material-models/hour of the same set of models run
  one at a time (the reference's experiment loop) and as one device batch.
  Both count, per model, the Ca/Cb upload, all time steps and the far field
  plus sigma0 from the device phasor sums; the incident beams are generated
  beforehand (cached per beam key) and reported separately.
* ``serving``: a Poisson stream of requests through the continuous-batching
  server with an HBM budget; p50/p95/p99 latency and models/hour.
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


def batched_vs_plain(count=8, chunk=64, device=0, log=print):
    native.init(device)
    t0 = time.perf_counter()
    cases = batched_workload(count)
    beams = B.BeamCache()
    for pc, _, _, _ in cases:
        beams.warm(pc)
    t_beams = time.perf_counter() - t0
    cells = [pc.geom_params.size_x * pc.geom_params.size_y * pc.geom_params.size_z
             for pc, _, _, _ in cases]
    steps = [pc.geom_params.size_t for pc, _, _, _ in cases]
    log(f'batched_vs_plain: {len(cases)} models, {len(beams)} beams in {t_beams:.1f} s')

    # plain: one model at a time
    models = []
    t0 = time.perf_counter()
    plain_back = []
    for pc, Ca, Cb, name in cases:
        m = B.DeviceModel(pc, Ca, Cb, beams, name)
        B.run_plain([m])
        plain_back.append(m.sigma0()['sigma_h'])
        m.release()
    t_plain = time.perf_counter() - t0
    log(f'batched_vs_plain: plain {t_plain:.2f} s')

    # batched: all resident, one launch per kernel per step
    t0 = time.perf_counter()
    models = _models_for(cases, beams)
    B.run_batched(models, chunk=chunk)
    batch_back = [m.sigma0()['sigma_h'] for m in models]
    t_batch = time.perf_counter() - t0
    for m in models:
        m.release()
    beams.release()
    same = all(np.array_equal(a, b) for a, b in zip(plain_back, batch_back))
    log(f'batched_vs_plain: batched {t_batch:.2f} s, identical results: {same}')
    n = len(cases)
    cu = float(np.dot(cells, steps))
    return {'models': n, 'beams': len(set(map(B.beam_key, [c[0] for c in cases]))),
            'beam_setup_s': round(t_beams, 2),
            'plain_s': round(t_plain, 3), 'batched_s': round(t_batch, 3),
            'plain_models_per_hour': round(n / t_plain * 3600, 1),
            'batched_models_per_hour': round(n / t_batch * 3600, 1),
            'speedup': round(t_plain / t_batch, 2),
            'batched_Mcell_updates_per_s': round(cu / t_batch / 1e6, 1),
            'plain_Mcell_updates_per_s': round(cu / t_plain / 1e6, 1),
            'identical_results': bool(same),
            'grid_cells_min_max': [int(min(cells)), int(max(cells))],
            'steps': int(max(steps))}


def serving_requests(n, rate_hz, seed=7, lateral=(4.0, 6.0), height=(2.0, 3.0),
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


def serving(n=24, rate_hz=None, budget_gb=None, max_slots=16, chunk=96, device=0, log=print,
            device_materials=False, use_graph=True):
    """Serve ``n`` Poisson requests; the rate defaults to 1.5x the service
    rate of the same requests run one at a time, so admission binds."""
    native.init(device)
    reqs = serving_requests(n, rate_hz=1.0, device_materials=device_materials)
    be = HipBackend(device, use_graph=use_graph)
    for r in reqs:                         # beams first: reported separately
        be.beams.warm(r.params)
    fps = [be.footprint(r) for r in reqs]
    # the same requests run one at a time (admit, run, sigma0, release) give
    # the plain service rate, which sets the arrival rate
    svc = None
    if rate_hz is None:
        t0 = time.perf_counter()
        k = n
        for r in reqs[:k]:
            m = be.admit(r)
            B.run_plain([m])
            m.sigma0()
            m.release()
        svc = (time.perf_counter() - t0) / k
        rate_hz = 1.5 / svc
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    t = 0.0
    for r, gap in zip(reqs, gaps):       # rescale the Poisson stream to the rate
        t += gap / rate_hz
        r.arrival = t
    budget = int((budget_gb or 0) * (1 << 30)) or int(4 * np.median(fps))
    srv = Server(be, hbm_budget=budget, max_slots=max_slots, chunk_steps=chunk)
    t0 = time.perf_counter()
    res = srv.serve(reqs)
    wall = time.perf_counter() - t0
    be.beams.release()
    st = latency_stats(res)
    st.update({'requests': n, 'errors': sum(r.error is not None for r in res),
               'materials': 'built on the device' if device_materials else 'host volumes uploaded',
               'arrival_rate_hz': round(rate_hz, 3), 'hbm_budget_bytes': budget,
               'peak_hbm_bytes': srv.peak_used, 'peak_members': srv.peak_slots,
               'chunks': srv.chunks, 'wall_s': round(wall, 2),
               'boundary_host_s': {k: round(v, 3) for k, v in srv.phase_s.items()}})
    if svc:
        # the same requests run one at a time (admit, run, sigma0, release)
        st['plain_models_per_hour'] = round(3600.0 / svc, 1)
        if st.get('models_per_hour'):
            st['served_vs_plain'] = round(st['models_per_hour'] * svc / 3600.0, 3)
    log(f'serving: {st}')
    return st


def batch_step_rate(count=8, steps=30, launch=None, device=0):
    """Steady-state step rate of the batched workload as one batch: all
    members resident, ``steps`` graph-replayed steps after a short warm-up.
    Returns ms per step, Gcell-updates/s and the algorithmic rate (the
    split kernels' and the DFT's compulsory bytes per step over the step)."""
    from .roofline import algorithmic_bytes
    native.init(device)
    cases = batched_workload(count)
    beams = B.BeamCache()
    models = _models_for(cases, beams)
    alg, cells = 0, 0
    for pc, _, _, _ in cases:
        g = pc.geom_params
        a = algorithmic_bytes(g.size_x, g.size_y, g.size_z, g.size_pml, None)
        alg += a['h'] + a['e'] + a['dft']
        cells += g.size_x * g.size_y * g.size_z
    b = native.Batch([m.sim for m in models])
    if launch:
        b.set_launch(*launch)
    b.run(5, False)
    b.sync()
    ms = b.time_steps(steps, True) / steps
    b.destroy()
    for m in models:
        m.release()
    beams.release()
    return {'models': len(cases), 'cells': cells, 'steps': steps, 'ms_per_step': round(ms, 4),
            'Gcell_updates_per_s': round(cells / ms / 1e6, 1),
            'alg_TBs': round(alg / ms / 1e9, 2)}


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
        out['batch_step'] = batch_step_rate()
    print(json.dumps(out))
