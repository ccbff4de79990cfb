"""Tile-shape selection for the stencil kernels.

A launch shape is (ty, zchunk, xvec, xwaves, xcd_order).  Each lane owns
xvec consecutive x cells (1, 2 or 4; 0 = chosen by the library from the
grid width) and marches z over zchunk cells (0 = whole column).  Split
stepper: blocks of xwaves x ty wavefronts (at most 4), xwaves side by side
along x on each of ty rows.  Fused stepper: ty in {2, 4, 8} rows walked by
one wavefront, xwaves 1.  Halo stepper: the split stepper's blocks, 1 or 2
cells per lane.  xwaves -1 (split stepper, ty 0, one cell per lane): flat
tiles, each wavefront over 64 consecutive cells of a plane's row-major
(y, x) index, so that narrow rows share wavefronts.  xcd_order 1 deals each XCD a contiguous range of
tiles, 0 runs them in plain order, 2 gives each XCD a band of rows per z
chunk and 3 + b (split and halo steppers) deals the XCDs bands of 2^b row
blocks in turn.
``autotune`` measures candidates on the device and records the best per
(stepper, arch, grid) in a JSON cache; ``launch_for`` reads it, falling back
to a default.
"""
import json
import os

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tuning_cache.json')
CANDIDATES = {
    'fused': [(r, zc, xv, 1, xo) for xv, rows in ((1, (2, 4, 8)), (2, (2, 4)), (4, (2,)))
              for r in rows for zc in (4, 8, 16, 32) for xo in (1, 0)],
    'split': [(ty, zc, xv, 1, 1) for xv in (4, 2, 1) for ty in (1, 2, 4) for zc in (8, 16, 32)]
             + [(ty, zc, xv, xw, xo) for xv in (1, 2) for xw, ty in ((4, 1), (2, 1), (2, 2))
                for zc in (1, 2, 3, 4, 6, 8, 16) for xo in (0, 1)]
             + [(0, zc, 1, -1, xo) for zc in (1, 2, 3, 4, 6) for xo in (0, 4, 5)],
    'rows': [(r, zc, xv, 1, xo) for r in (3, 7) for zc in (4, 8, 16, 32) for xv in (1, 2)
             for xo in (0, 1)],
    'halo': [(ty, zc, xv, xw, 0) for xv in (1, 2) for xw, ty in ((4, 1), (2, 1), (2, 2), (1, 4))
             for zc in (4, 8, 16)]
            # fine bands of 1, 2 and 4 row blocks per XCD (the wide grids'
            # neighbour rows then pass through one L2)
            + [(ty, zc, 1, xw, xo) for xw, ty in ((2, 1), (4, 1), (2, 2)) for zc in (8, 16)
               for xo in (3, 4, 5)],
}
# library defaults: 0 = chosen from the grid (split: one cell per lane,
# four wavefronts per block, as many along x as the grid fills); halo: one
# row of two wavefronts per block, XCDs dealt bands of 4 row blocks (at
# 761 x 761 x 203 and 1523 x 1523 x 311 no slower than 2 x 2 blocks in plain
# order, and 5 % faster at the larger)
DEFAULT = {'fused': (4, 8, 0, 1, 0), 'split': (0, 2, 0, 0, 0), 'rows': (7, 16, 2, 1, 0),
           'halo': (1, 8, 1, 2, 5)}
STEPPERS = {0: 'split', 1: 'fused', 2: 'rows', 3: 'halo'}


def _stepper(fused):
    return STEPPERS[int(fused)]


def _shape(hit):
    return (int(hit['ty']), int(hit['zchunk']), int(hit.get('xvec', 0)),
            int(hit.get('xwaves', 1)), int(hit.get('xcd_order', 1)))


def _key(nx, ny, nz, fused, arch='gfx950'):
    return f'{_stepper(fused)}:{arch}:{nx}x{ny}x{nz}'


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def launch_for(nx, ny, nz, fused=True, arch='gfx950'):
    """Cached choice for this grid, else the stepper's default."""
    hit = _load().get(_key(nx, ny, nz, fused, arch))
    if hit:
        return _shape(hit)
    return DEFAULT[_stepper(fused)]


def autotune(sim, nx, ny, nz, fused=True, arch='gfx950', steps=10, candidates=None, save=True):
    """Time each candidate shape on ``sim`` (a ready native.Sim whose stepper
    is already chosen) and keep the fastest.  Returns (shape,
    {shape: ms_per_step})."""
    from .._native import FdtdError
    results = {}
    # the stepper's default shape is always measured, so a cached shape is
    # one that beat it on this grid
    shapes = list(candidates or CANDIDATES[_stepper(fused)])
    default = tuple(DEFAULT[_stepper(fused)])
    if default not in shapes:
        shapes.insert(0, default)
    for shape in shapes:
        try:
            sim.set_launch(*shape)
        except FdtdError:
            continue                       # shape not valid for this grid / stepper
        sim.reset()
        sim.run(2)
        sim.sync()
        ms = sim.time_steps(steps, per_kernel=False)['total'] / steps
        results[tuple(shape)] = ms
    best = min(results, key=results.get)
    sim.set_launch(*best)
    if save:
        data = _load()
        data[_key(nx, ny, nz, fused, arch)] = {'ty': best[0], 'zchunk': best[1],
                                               'xvec': best[2], 'xwaves': best[3],
                                               'xcd_order': best[4], 'ms_per_step': results[best]}
        tmp = CACHE + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    return best, results


#: Execution candidates for a batch: (label, DeviceBatch mode argument).
#: 'auto' is the rule of hip.batch.auto_mode; the others fix the mode.
BATCH_CANDIDATES = [
    ('packed-z2', ('packed', (0, 2, 0, 0, 0))),
    ('packed-z4', ('packed', (0, 4, 0, 0, 0))),
    ('serial-1', ((1, 0, 1), None)),
    ('serial-2', ((1, 0, 2), None)),
    ('serial-4', ((1, 0, 4), None)),
    ('serial-4-small-first', ((1, 8_000_000, 4), None)),
]


def batch_key(models, arch='gfx950'):
    """Cache key of a batch: its members' grids, sorted."""
    sizes = sorted(tuple(m.sim.shape) for m in models)
    return f'batch:{arch}:' + ';'.join('x'.join(map(str, sz[::-1])) for sz in sizes)


def batch_mode_for(models, arch='gfx950'):
    """Cached execution of this batch (a (mode, launch) pair for
    DeviceBatch), else None."""
    hit = _load().get(batch_key(models, arch))
    if not hit:
        return None
    launch = tuple(hit['launch']) if hit.get('launch') else None
    mode = hit['mode'] if isinstance(hit['mode'], str) else tuple(hit['mode'])
    return mode, launch


def autotune_batch(models, steps=8, arch='gfx950', candidates=None, save=True):
    """Time each execution candidate on the batch of ``models`` (ready
    DeviceModels, before their run: their state is reset afterwards, so the
    results do not depend on the tuning) and keep the fastest.  Returns
    ((mode, launch), {label: ms_per_step})."""
    import time
    from .batch import DeviceBatch
    results, choice = {}, {}
    for label, (mode, launch) in candidates or BATCH_CANDIDATES:
        db = DeviceBatch(models, launch=launch, mode=mode)
        db.run(2)
        db.sync()
        t0 = time.perf_counter()
        db.run(steps)
        db.sync()
        results[label] = (time.perf_counter() - t0) * 1e3 / steps
        choice[label] = (mode, launch)
        db.destroy()
    for m in models:
        m.sim.reset()
        m.steps_done = 0
    best = min(results, key=results.get)
    if save:
        mode, launch = choice[best]
        data = _load()
        data[batch_key(models, arch)] = {'label': best, 'mode': mode, 'launch': launch,
                                         'ms_per_step': results[best]}
        tmp = CACHE + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    return choice[best], results


def _workload_grids(name):
    """(params, Ca, Cb) of each distinct grid of the batched workload (32
    soil realizations) or the serving workload (the canopy requests' lateral
    sizes x heights), on their own materials."""
    from .. import workloads
    from ..materials import compute_material_arrays, synthetic
    from . import measure
    out = {}
    if name == 'batched':
        for pc, Ca, Cb, _ in measure.batched_workload(32):
            g = pc.geom_params
            out.setdefault((g.size_x, g.size_y, g.size_z), (pc, Ca, Cb))
        return list(out.values())
    for lat in measure.CANOPY_LATERAL:
        for h in measure.CANOPY_HEIGHT:
            pc = workloads.base_params(incident_duration=10.0, illumination_width=2.0,
                                       simulation_duration=12.0, material_fraction=0.6)
            workloads.set_geom_params(pc, lat, lat, h, 12.0, 0.025, 'wavelength')
            g = pc.geom_params
            m = synthetic.canopy_over_soil(g.size_x, g.size_y, g.size_z_material, g.cell_size,
                                           workloads.lam_of(pc), max(2, g.size_z_material // 3),
                                           seed=1, density=4.0)
            Ca, Cb = compute_material_arrays(pc, m)
            out[(g.size_x, g.size_y, g.size_z)] = (pc, Ca, Cb)
    return list(out.values())


def _tune_grids(name, fused):
    """Autotune each distinct grid of the batched or serving workload on its
    own materials; returns {grid: (best, results)}."""
    from .. import _native as native, workloads
    from . import common
    native.init(0)
    found = {}
    for pc, Ca, Cb in _workload_grids(name):
        g = pc.geom_params
        sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
        sim.set_stepper(fused)
        beam = native.Beam(g.size_t_inc, g.size_y, g.size_x)
        beam.upload(0, *workloads.synthetic_beam_planes(pc, 0, min(16, g.size_t_inc)))
        sim.attach_beam(beam)
        sim.set_materials(Ca, Cb)
        found[(g.size_x, g.size_y, g.size_z)] = autotune(sim, g.size_x, g.size_y, g.size_z, fused=fused, steps=20)
        sim.destroy()
        beam.release()
    return found


def _tune_workload(name, fused):
    """Build a ready simulation of a named workload with uniform-value inputs
    (timing does not depend on the values) and autotune it."""
    import numpy as np
    from .. import _native as native, workloads
    from . import common
    pc = {'large': workloads.large_params, 'plain': workloads.plain_params,
          'largest': workloads.largest_params}[name]()
    g = pc.geom_params
    native.init(0)
    sim = native.Sim(common.main_geom(pc), common.coeffs_for(pc))
    sim.set_stepper(fused)
    beam = native.Beam(g.size_t_inc, g.size_y, g.size_x)
    beam.upload(0, *workloads.synthetic_beam_planes(pc, 0, 16))
    sim.attach_beam(beam)
    shape = (g.size_z, g.size_y, g.size_x)
    # vacuum over a lower 70 % whose planes hold other values, so the E
    # kernel reads Ca/Cb where a ground model would (its timing depends on
    # that, not on the values)
    ca = np.ones(shape, np.float32)
    cb = np.full(shape, common.update_scalars(pc)['dte0dx'], np.float32)
    ca[:int(0.7 * g.size_z)] = np.float32(0.99)
    sim.set_materials(ca, cb)
    best, res = autotune(sim, g.size_x, g.size_y, g.size_z, fused=fused)
    sim.destroy()
    beam.release()
    return best, res


if __name__ == '__main__':
    # python -m sar_pixel_simulator_amd.hip.tuning [workload ...] [--steppers 3,0]
    # workloads: large, plain, largest (one grid each), batched, serving
    # (each of their distinct grids)
    import sys
    args = sys.argv[1:]
    steppers = (1, 0, 2, 3)
    if '--steppers' in args:
        i = args.index('--steppers')
        steppers = tuple(int(v) for v in args[i + 1].split(','))
        del args[i:i + 2]
    for name in args or ['large']:
        for fused in steppers:
            if name in ('batched', 'serving'):
                # every distinct grid of the workload (the split stepper's sizes)
                for grid, (best, res) in _tune_grids(name, fused).items():
                    default = res[tuple(DEFAULT[_stepper(fused)])]
                    print(f'{name} {_stepper(fused)} {grid}: best {best} {res[best]:.4f} ms/step '
                          f'(default shape {default:.4f})', flush=True)
                continue
            best, res = _tune_workload(name, fused)
            tag = _stepper(fused)
            for shape, ms in sorted(res.items(), key=lambda kv: kv[1]):
                print(f'{name} {tag} ty={shape[0]} zchunk={shape[1]} xvec={shape[2]} '
                      f'xwaves={shape[3]} xcd_order={shape[4]}: {ms:.4f} ms/step', flush=True)
            print(f'{name} {tag}: best {best}', flush=True)
