"""Throughput of the FDTD time loop on one MI355X.

Workload: the single large plain run at the README shape -- 16 x 16 x 3
lambda at 0.021 lambda (761 x 761 x 203 cells = 1.18e8, 1714 steps at
Courant 0.5), 5.4 GHz HH theta 20, a 1 lambda snow layer of Bernoulli(0.3)
ice voxels over rough soil.  One step = one full Yee time step (H half with
CPML and E_inc injection, E half with CPML and H_inc injection, phasor DFT on
the six planes) of that domain.  Materials, CPML state, phasor sums and the
incident beam are resident in HBM before the timed region; the beam is an
analytic Gaussian-tapered plane wave (synthetic) of the real beam's shape.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

Prints one JSON line (see the keys below).
"""
import argparse
import glob
import json
import os
import shutil
import sqlite3
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from sar_pixel_simulator_amd.hip.roofline import algorithmic_bytes, cpml_cells  # noqa: E402

METRIC = 'FDTD Mcell-updates/s + material-models/hr batched vs plain; p50/p99 latency'
HBM_PEAK_GBS = 8000.0
PROBE_BYTES = 1 << 30          # calibration copy: 1 GiB read + 1 GiB written
LAUNCH_KEYS = ('ty', 'zchunk', 'xvec', 'xwaves', 'xcd_order')
STEPPER_IDS = {'split': 0, 'fused': 1, 'rows': 2, 'halo': 3}
PMC_COUNTERS = ('FETCH_SIZE', 'WRITE_SIZE')


def uniform_top(Ca, Cb):
    """First z of the run of top planes whose cells all hold one (Ca, Cb)
    pair (the vacuum above the material); the E kernel reads no Ca/Cb there."""
    nz = Ca.shape[0]
    a0, b0 = Ca[-1].flat[0], Cb[-1].flat[0]
    z = nz
    while z > 0 and np.all(Ca[z - 1] == a0) and np.all(Cb[z - 1] == b0):
        z -= 1
    return z


#: Host cores a one-GPU job may use on the MI355X box (its CPU share for
#: one GPU); the CPU baseline uses all of them, or fewer if fewer are visible.
HOST_CORE_SHARE = 16


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, HOST_CORE_SHARE))


def cfg0_end_to_end(log=print):
    """The plain CPU-sized workload (81 x 81 x 81 cells, 480 steps, rough
    soil, HH theta 20) end to end, one model: on the host through the oracle
    (tapered-wave excitation, incident FDTD, main FDTD, top-face far field
    and sigma0, float32 fields, OpenMP on the host cores) and on the device
    (beam synthesis, main run, far field and sigma0)."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    import beam_oracle
    import helpers as H
    import oracle
    from sar_pixel_simulator_amd import _native as native, workloads
    from sar_pixel_simulator_amd.hip import batch as B
    from sar_pixel_simulator_amd.materials import compute_material_arrays
    oracle.build()
    pc = workloads.plain_params()
    m = workloads.plain_material(pc, seed=0)
    g = pc.geom_params
    nthreads = cpu_threads()
    t0 = time.perf_counter()
    pex, pey = beam_oracle.create_inc_excitation(pc)
    rec, inc = H.oracle_incident(pc, pex, pey, nthreads=nthreads)
    Ca, Cb = compute_material_arrays(pc, m)
    sim = H.oracle_main(pc, Ca, Cb, rec, nthreads=nthreads)
    n = g.size_pml
    nxp, nyp, nzp = g.size_x - 2 * n, g.size_y - 2 * n, g.size_z - 2 * n
    th = np.linspace(-80.0, 80.0, 161)
    ff = oracle.far_field(sim.phasors(), nxp, nyp, nzp, g.cell_size, 2000.0, th, [0.0],
                          pc.inc_beam_params.frequency)
    oracle.sigma0(ff, oracle.incident_power(inc, g.cell_size), nxp, nyp, g.cell_size)
    t_cpu = time.perf_counter() - t0
    native.init(0)
    t0 = time.perf_counter()
    beams = B.BeamCache()
    dm = B.DeviceModel.from_material(pc, m, beams)
    B.run_plain([dm])
    dm.sigma0()
    dm.release()
    beams.release()
    t_gpu = time.perf_counter() - t0
    out = {'workload': f'plain CPU-sized run {g.size_x}x{g.size_y}x{g.size_z}, {g.size_t} steps, '
                       'one model end to end (beam, incident run, main run, far field, sigma0)',
           'cpu_s': round(t_cpu, 3), 'cpu_cores': nthreads, 'gpu_s': round(t_gpu, 3),
           'gpu_vs_cpu': round(t_cpu / t_gpu, 2)}
    log(f'bench: cfg0 end to end {out}')
    return out


def cpu_baseline(pc, Ca, Cb, target_s=15.0):
    """The oracle (C restatement of the reference's CPU loops, OpenMP over
    z slabs as cpu/fdtd_sim.py does) timed on a bounded sample of the same
    workload: a few steps of the full domain with injection active."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    import oracle
    import helpers as H
    from sar_pixel_simulator_amd import workloads
    oracle.build()
    g = pc.geom_params
    nthreads = cpu_threads()
    max_steps = 160
    beam = workloads.synthetic_beam_planes(pc, 0, max_steps)
    geo = H.oracle_geom(pc)
    geo['nt_inc'] = max_steps
    sim = oracle.OracleSim(geo, Ca, Cb, H.profiles_of(pc), inc=beam, nthreads=nthreads)
    t0 = time.perf_counter()
    sim.run(1)
    t1 = time.perf_counter() - t0
    n = int(max(2, min(max_steps - 1, target_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    sim.run(n)
    el = time.perf_counter() - t0
    cells = g.size_x * g.size_y * g.size_z
    return {'value': cells * n / el / 1e6, 'unit': 'Mcell-updates/s', 'cores': nthreads,
            'host_cpus_visible': os.cpu_count(), 'kind': 'port',
            'sample': f'{n} time steps (after 1 warm-up) of the full {g.size_x}x{g.size_y}x'
                      f'{g.size_z} domain with injection active, {el:.1f} s'}


def rocprofv3():
    for cand in (shutil.which('rocprofv3'), '/opt/rocm/bin/rocprofv3'):
        if cand and os.path.exists(cand):
            return cand
    return None


def pmc_pass(counter, child_args, out_dir):
    """One ``rocprofv3 --pmc <counter>`` run of a short child bench (a few
    steps, then the calibration copy).  Returns the path of its database."""
    d = os.path.join(out_dir, counter)
    cmd = ['timeout', '-s', 'KILL', '120', rocprofv3(), '--pmc', counter, '-d', d, '-o', 'pmc',
           '--', sys.executable, os.path.join(ROOT, 'bench.py'), '--pmc-child'] + child_args
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    dbs = glob.glob(os.path.join(d, '**', '*results.db'), recursive=True)
    if r.returncode != 0 or not dbs:
        tail = '\n'.join(r.stdout.splitlines()[-5:])
        raise RuntimeError(f'rocprofv3 --pmc {counter} exited {r.returncode}: {tail}')
    return dbs[0]


def counter_kib(db, counter, kernel):
    """Average value (KiB per dispatch) of ``counter`` over the dispatches
    whose kernel name contains ``kernel``."""
    c = sqlite3.connect(db)
    try:
        rows = [v for name, v in c.execute(
            'select kernel_name, value from counters_collection where counter_name = ?',
            (counter,)) if kernel in name]
    finally:
        c.close()
    if not rows:
        raise RuntimeError(f'no {kernel} dispatch with {counter} in {db}')
    return sum(rows) / len(rows)


def measure_traffic(kernel, child_args, keep_dir=None):
    """HBM bytes per launch of ``kernel`` from FETCH_SIZE and WRITE_SIZE,
    each scaled by what the same counter reads for the calibration copy of
    PROBE_BYTES (k_copy_probe<1>: PROBE_BYTES fetched, PROBE_BYTES written)."""
    if rocprofv3() is None:
        return None, {'note': 'rocprofv3 not found'}
    out_dir = keep_dir or tempfile.mkdtemp(prefix='fdtd_pmc_')
    detail = {}
    try:
        total = 0.0
        for ctr in PMC_COUNTERS:
            db = pmc_pass(ctr, child_args, out_dir)
            k_kib = counter_kib(db, ctr, 'fdtd::' + kernel)
            c_kib = counter_kib(db, ctr, 'fdtd::k_copy_probe<1>')
            scale = PROBE_BYTES / (c_kib * 1024.0)
            nbytes = k_kib * 1024.0 * scale
            total += nbytes
            key = ctr.split('_')[0].lower()
            detail[key + '_bytes'] = int(nbytes)
            detail[key + '_raw_kib'] = round(k_kib, 1)
            detail[key + '_scale'] = round(scale, 4)
        return int(total), detail
    except (RuntimeError, sqlite3.Error) as e:
        return None, {'note': str(e)[:300]}
    finally:
        if keep_dir is None:
            shutil.rmtree(out_dir, ignore_errors=True)


def largest_leg(fused, steps=10, warmup=3, log=print):
    """Throughput at the largest single domain (32 x 32 x 4 lambda at 0.021
    lambda: 1523 x 1523 x 311 = 7.2e8 cells, 1.5 lambda of snow over rough
    soil, VV theta 30), everything resident in HBM (the 50.9 GB beam buffer
    included; its first planes hold an analytic beam of the real shape).
    The materials are drawn on the device (same statistics as the host
    generator; the step rate does not depend on the values)."""
    from sar_pixel_simulator_amd import _native as native, workloads
    from sar_pixel_simulator_amd.hip import common, tuning
    from sar_pixel_simulator_amd.materials import device
    pc = workloads.largest_params()
    g = pc.geom_params
    t0 = time.perf_counter()
    snow = int(round(1.5 / 0.021))
    spec = device.snow_over_soil_spec(pc, g.size_z_material - snow, snow, seed=0)
    geom = common.main_geom(pc)
    sim = native.Sim(geom, common.coeffs_for(pc))
    sim.set_stepper(fused)
    shape = tuning.launch_for(g.size_x, g.size_y, g.size_z, fused=fused)
    sim.set_launch(*shape)
    beam = native.Beam(g.size_t_inc, g.size_y, g.size_x)
    beam.upload(0, *workloads.synthetic_beam_planes(pc, 0, warmup + 2 * steps + 2))
    sim.attach_beam(beam)
    sim.build_materials(*device.to_c(spec, pc))
    sim.sync()
    log(f'bench: largest {g.size_x}x{g.size_y}x{g.size_z} state resident in {time.perf_counter() - t0:.1f} s')
    uni_z = g.size_z_material          # vacuum (uniform Ca, Cb) above the material
    sim.run(warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    k = sim.time_steps(steps, per_kernel=True)
    cells = g.size_x * g.size_y * g.size_z
    ab = algorithmic_bytes(g.size_x, g.size_y, g.size_z, g.size_pml, uni_z, taps=fused in (0, 3))
    step_bytes = (ab['h'] + ab['e'] - 24 * cells if fused else ab['h'] + ab['e']) + ab['dft']
    ms = wall * 1e3 / steps
    out = {'workload': f'largest single domain: {g.size_x}x{g.size_y}x{g.size_z} cells (32x32x4 '
                       'lambda at 0.021 lambda), 1.5 lambda of snow over rough soil, VV theta 30, '
                       f'{g.size_t}-step run',
           'cells': cells, 'steps': steps, 'ms_per_step': round(ms, 4),
           'Mcell_updates_per_s': round(cells / (ms * 1e-3) / 1e6, 1),
           'step_alg_GBs': round(step_bytes / (ms * 1e-3) / 1e9, 1),
           'kernels_ms': {kk: round(v, 4) for kk, v in k.items()},
           'kernels_alg_GBs': ({kk: round(ab[kk] / (k[kk] * 1e-3) / 1e9, 1) for kk in ('h', 'e', 'dft')}
                               if not fused and min(k['h'], k['e'], k['dft']) > 0 else
                               {'fused': round((ab['h'] + ab['e'] - 24 * cells) / (k['h'] * 1e-3) / 1e9, 1),
                                'dft': round(ab['dft'] / (k['dft'] * 1e-3) / 1e9, 1)}
                               if fused and min(k['h'], k['dft']) > 0 else None),
           'stepper': tuning.STEPPERS[int(fused)],
           'launch': dict(zip(LAUNCH_KEYS, shape)),
           'device_bytes': int(native.footprint(geom) + 16 * g.size_t_inc * g.size_y * g.size_x)}
    sim.destroy()
    beam.release()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-extras', action='store_true',
                    help='skip the batched-vs-plain and serving legs')
    ap.add_argument('--stepper', choices=['split', 'fused', 'rows', 'halo'], default=None,
                    help='time-step kernels (default: the library default)')
    ap.add_argument('--ty', type=int, default=0)
    ap.add_argument('--zchunk', type=int, default=-1)
    ap.add_argument('--xvec', type=int, default=-1,
                    help='x cells per lane of the split kernels (0 = automatic)')
    ap.add_argument('--xwaves', type=int, default=0,
                    help='split stepper: wavefronts side by side along x per block')
    ap.add_argument('--xcd-order', type=int, default=-1,
                    help='1: XCD-contiguous tile order, 0: plain order')
    ap.add_argument('--no-traffic', action='store_true',
                    help='skip the two rocprofv3 --pmc passes that measure HBM traffic')
    ap.add_argument('--pmc-dir', default=None,
                    help='keep the PMC databases in this directory')
    ap.add_argument('--pmc-child', action='store_true', help=argparse.SUPPRESS)
    ap.add_argument('--batch-count', type=int, default=32,
                    help='batched leg: soil realizations (x HH/VV)')
    ap.add_argument('--serve-requests', type=int, default=200, help='serving legs: requests')
    ap.add_argument('--reference-loop', type=int, default=24,
                    help='serving leg: requests timed through the reference-style loop')
    ap.add_argument('--largest-only', action='store_true',
                    help='run only the largest-domain leg and print its JSON (for profiling it)')
    args = ap.parse_args()
    if args.pmc_child:
        args.no_cpu = args.no_extras = args.no_traffic = True
    if args.gpus != 1:
        print(json.dumps({'error': 'the engine runs on one GPU; independent models shard '
                                   'across GPUs as separate processes, with no device '
                                   'communication, so there is no multi-GPU run to measure',
                          'n_gpus': args.gpus}))
        sys.exit(2)

    from sar_pixel_simulator_amd import _native as native, workloads
    from sar_pixel_simulator_amd.hip import common, tuning
    from sar_pixel_simulator_amd.materials import compute_material_arrays

    if args.largest_only:
        native.init(0)
        lg = workloads.largest_params().geom_params
        fused = (common.stepper_for(lg.size_x, lg.size_y, lg.size_z) if args.stepper is None
                 else STEPPER_IDS[args.stepper])
        print(json.dumps(largest_leg(fused, steps=args.steps, warmup=args.warmup,
                                     log=lambda m: print(m, file=sys.stderr, flush=True))),
              flush=True)
        return

    pc = workloads.large_params()
    g = pc.geom_params
    print(f'bench: grid {g.size_x}x{g.size_y}x{g.size_z}, nt {g.size_t}, '
          f'nt_inc {g.size_t_inc}', file=sys.stderr, flush=True)
    mat = workloads.large_material(pc, seed=0)
    Ca, Cb = compute_material_arrays(pc, mat)
    del mat

    native.init(0)
    info = native.device_info()
    fused = (common.stepper_for(g.size_x, g.size_y, g.size_z) if args.stepper is None
             else STEPPER_IDS[args.stepper])
    stepper_name = tuning.STEPPERS[fused]
    fused_kernel = {1: 'k_step_strip', 2: 'k_step_rows', 3: 'k_step_halo'}.get(fused)
    sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
    sim.set_stepper(fused)
    ty, zc, xv, xw, xo = tuning.launch_for(g.size_x, g.size_y, g.size_z, fused=fused)
    if args.ty > 0:
        ty = args.ty
    if args.zchunk >= 0:
        zc = args.zchunk
    if args.xvec >= 0:
        xv = args.xvec
    if args.xwaves > 0:
        xw = args.xwaves
    if args.xcd_order >= 0:
        xo = args.xcd_order
    launch = (ty, zc, xv, xw, xo)
    sim.set_launch(*launch)
    beam = native.Beam(g.size_t_inc, g.size_y, g.size_x)
    n_fill = min(g.size_t_inc, args.warmup + args.steps + 8)
    beam.upload(0, *workloads.synthetic_beam_planes(pc, 0, n_fill))
    sim.attach_beam(beam)
    sim.set_materials(Ca, Cb)
    print('bench: device state resident; warm-up', file=sys.stderr, flush=True)

    # The library's HIP runtime is /opt/rocm's; torch ships its own, so the
    # timed region is bracketed by a synchronisation of the library stream
    # the steps run on (nothing else runs on the device).
    sim.run(args.warmup)
    sim.sync()
    if args.pmc_child:
        # profiled child: a few steps, then the calibration copy
        sim.run(args.steps)
        sim.sync()
        native.probe_copy(PROBE_BYTES, vec=1, reps=1)
        sim.destroy()
        beam.release()
        return
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    cells = g.size_x * g.size_y * g.size_z
    value = cells * args.steps / wall / 1e6
    print(f'bench: {args.steps} steps in {wall:.3f} s', file=sys.stderr, flush=True)

    # per-kernel HIP-event times on the library's stream
    k = sim.time_steps(min(args.steps, 50), per_kernel=True)
    uni_z = uniform_top(Ca, Cb)
    ab = algorithmic_bytes(g.size_x, g.size_y, g.size_z, g.size_pml, uni_z, taps=fused in (0, 3))

    def gbs(nbytes, ms):
        return round(nbytes / (ms * 1e-3) / 1e9, 1) if ms > 0 else None

    if fused:
        # one kernel does both halves: E and H each read once and written
        # once (the H-half reads of E and E-half reads of H are gone)
        fb = ab['h'] + ab['e'] - 24 * ab['cells']
        kern = {'name': fused_kernel, 'bytes': fb, 'ms': k['h']}
        per_kernel = {fused_kernel + '_ms': round(k['h'], 4), 'k_dft_ms': round(k['dft'], 4),
                      'fused_GBs': gbs(fb, k['h']), 'dft_GBs': gbs(ab['dft'], k['dft'])}
        step_alg = fb + ab['dft']
    else:
        dom = 'e' if k['e'] >= k['h'] else 'h'
        kern = {'name': 'k_e_half' if dom == 'e' else 'k_h_half', 'bytes': ab[dom], 'ms': k[dom]}
        per_kernel = {'k_h_half_ms': round(k['h'], 4), 'k_e_half_ms': round(k['e'], 4),
                      'k_dft_ms': round(k['dft'], 4), 'h_GBs': gbs(ab['h'], k['h']),
                      'e_GBs': gbs(ab['e'], k['e']), 'dft_GBs': gbs(ab['dft'], k['dft'])}
        step_alg = ab['h'] + ab['e'] + ab['dft']
    per_kernel['step_alg_GBs'] = gbs(step_alg, ms_step)
    achieved = kern['bytes'] / (kern['ms'] * 1e-3) / 1e9
    sim.destroy()
    split_ref = None
    if fused and not args.no_extras:
        # the same workload on the split stepper, for comparison in the same
        # run: step time and its kernels against their algorithmic bytes
        print('bench: split stepper on the same workload', file=sys.stderr, flush=True)
        ssim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
        ssim.set_stepper(0)
        sshape = tuning.launch_for(g.size_x, g.size_y, g.size_z, fused=0)
        ssim.set_launch(*sshape)
        ssim.attach_beam(beam)
        ssim.set_materials(Ca, Cb)
        ssim.run(args.warmup)
        ssim.sync()
        t0 = time.perf_counter()
        ssim.run(args.steps)
        ssim.sync()
        s_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        sk = ssim.time_steps(min(args.steps, 50), per_kernel=True)
        sab = algorithmic_bytes(g.size_x, g.size_y, g.size_z, g.size_pml, uni_z, taps=True)
        split_ref = {'ms_per_step': round(s_ms, 4), 'Mcell_updates_per_s': round(cells / (s_ms * 1e-3) / 1e6, 1),
                     'launch': dict(zip(LAUNCH_KEYS, sshape)),
                     'k_h_half_ms': round(sk['h'], 4), 'k_e_half_ms': round(sk['e'], 4),
                     'k_dft_ms': round(sk['dft'], 4),
                     'k_e_half_frac': round(sab['e'] / (sk['e'] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     'k_h_half_frac': round(sab['h'] / (sk['h'] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     'step_alg_bytes': sab['h'] + sab['e'] + sab['dft'],
                     'split_step_over_this_step': round(s_ms / ms_step, 4)}
        ssim.destroy()
    beam.release()
    # practical ceiling: a plain streaming copy with the same 4-byte accesses
    # (and with 16-byte accesses)
    copy1_ms, copy1 = native.probe_copy(PROBE_BYTES, vec=1, reps=10)
    copy4_ms, copy4 = native.probe_copy(PROBE_BYTES, vec=4, reps=10)
    # the same access mix as the kernel (reads + writes per cell), streamed
    # flat: the practical ceiling for that mix
    mix = (8, 6) if fused else ((8, 3) if kern['name'] == 'k_e_half' else (6, 3))
    mix_ms, mix_gbs = native.probe_streams(cells, *mix, reps=5)
    # the kernel's own access pattern (z march, neighbour rows; for the halo
    # stepper the row below re-read and both buffer sets) without its
    # arithmetic: the practical ceiling of that pattern
    stencil = None
    if fused in (0, 3):
        kind = 2 if fused == 3 else 1 if kern['name'] == 'k_e_half' else 0
        st_ms, st_gbs = native.probe_stencil(g.size_x, g.size_y, g.size_z, zc if zc > 0 else g.size_z,
                                             kind, reps=5)
        stencil = {'half': ('H', 'E', 'fused step')[kind], 'zchunk': zc, 'GBs': round(st_gbs, 1)}
    traffic, tdetail = None, {}
    if not args.no_traffic:
        print('bench: PMC traffic passes (FETCH_SIZE, WRITE_SIZE)', file=sys.stderr, flush=True)
        child = ['--steps', '4', '--warmup', '1', '--stepper', stepper_name,
                 '--ty', str(ty), '--zchunk', str(zc), '--xvec', str(xv),
                 '--xwaves', str(xw), '--xcd-order', str(xo)]
        traffic, tdetail = measure_traffic(kern['name'], child, args.pmc_dir)
    roof = {'bound': 'hbm', 'kernel': kern['name'], 'achieved': round(achieved, 1),
            'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': traffic, 'alg_bytes_per_launch': kern['bytes'],
            'avg_launch_ms': round(kern['ms'], 4),
            'traffic_detail': tdetail,
            'copy_probe_GBs': {'4B_per_lane': round(copy1, 1), '16B_per_lane': round(copy4, 1)},
            'stream_probe': {'reads': mix[0], 'writes': mix[1], 'GBs': round(mix_gbs, 1)},
            'frac_of_stream_probe': round(achieved / mix_gbs, 4)}
    if stencil:
        roof['stencil_probe'] = stencil
        roof['frac_of_stencil_probe'] = round(achieved / stencil['GBs'], 4)
    if traffic:
        roof['traffic_GBs'] = round(traffic / (kern['ms'] * 1e-3) / 1e9, 1)

    cpu = None
    if not args.no_cpu:
        print('bench: CPU baseline sample', file=sys.stderr, flush=True)
        cpu = cpu_baseline(pc, Ca, Cb)
        cpu['plain_end_to_end'] = cfg0_end_to_end(
            log=lambda m: print(m, file=sys.stderr, flush=True))

    extras = {}
    if not args.no_extras:
        from sar_pixel_simulator_amd.hip import measure

        def log(msg):
            print(msg, file=sys.stderr, flush=True)
        print('bench: batched vs plain (64 uneven soil models)', file=sys.stderr, flush=True)
        extras['batched_vs_plain'] = measure.batched_vs_plain(count=args.batch_count, log=log)
        print('bench: serving, materials built on the device', file=sys.stderr, flush=True)
        extras['serving_device_materials'] = measure.serving(
            n=args.serve_requests, log=log, device_materials=True, reference_loop=args.reference_loop)
        print('bench: serving, host material volumes', file=sys.stderr, flush=True)
        extras['serving'] = measure.serving(n=args.serve_requests, log=log)
        print('bench: material setup, host vs device', file=sys.stderr, flush=True)
        extras['material_setup'] = measure.material_setup(log=log)
        print('bench: largest single domain', file=sys.stderr, flush=True)
        lg = workloads.largest_params().geom_params
        extras['largest'] = largest_leg(common.stepper_for(lg.size_x, lg.size_y, lg.size_z), log=log)

    out = {'metric': METRIC, 'value': round(value, 1), 'unit': 'Mcell-updates/s',
           'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
           'ms_per_step': round(ms_step, 4), 'higher_is_better': True, 'scaling': 'weak',
           'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
           'config': {'workload': 'single large plain run: 761x761x203 cells (16x16x3 '
                                  'lambda at 0.021 lambda, README shape), snow over rough '
                                  'soil, 5.4 GHz HH theta 20, 1714-step run',
                      'cells': cells, 'stepper': stepper_name,
                      'launch': dict(zip(LAUNCH_KEYS, launch)),
                      'uniform_planes_from_z': int(uni_z),
                      'device': info['name'], 'arch': info['arch'],
                      'cu_count': info['cu_count']},
           'roofline': roof, 'kernels': per_kernel, 'cpu_baseline': cpu}
    if split_ref:
        out['split_stepper_same_workload'] = split_ref
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
