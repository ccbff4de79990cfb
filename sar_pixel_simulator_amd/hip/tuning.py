"""Tile-shape selection for the fused stencil kernels.

A launch shape is (ty, zchunk): blocks of 64 x ty threads, each owning a
64 x ty column tile, marching zchunk cells in z (0 = whole column).  The
table below holds measured choices per shape class; ``autotune`` measures
candidates on the device and records the best in a JSON cache keyed by
device arch and grid dims.
"""
import json
import os

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tuning_cache.json')
CANDIDATES = [(1, 0), (2, 0), (4, 0), (1, 64), (2, 64), (4, 64), (4, 32), (2, 32)]


def _key(nx, ny, nz, arch='gfx950'):
    return f'{arch}:{nx}x{ny}x{nz}'


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def launch_for(nx, ny, nz, arch='gfx950'):
    """Cached choice for this grid, else the default (ty=4, whole column)."""
    hit = _load().get(_key(nx, ny, nz, arch))
    if hit:
        return int(hit['ty']), int(hit['zchunk'])
    return 4, 0


def autotune(sim, nx, ny, nz, arch='gfx950', steps=10, candidates=CANDIDATES, save=True):
    """Time each candidate shape on ``sim`` (a ready native.Sim) and keep the
    fastest.  Returns ((ty, zchunk), {shape: ms_per_step})."""
    results = {}
    for ty, zc in candidates:
        sim.set_launch(ty, zc)
        sim.reset()
        sim.run(2)
        sim.sync()
        ms = sim.time_steps(steps, per_kernel=False)['total'] / steps
        results[(ty, zc)] = ms
    best = min(results, key=results.get)
    sim.set_launch(*best)
    if save:
        data = _load()
        data[_key(nx, ny, nz, arch)] = {'ty': best[0], 'zchunk': best[1],
                                        'ms_per_step': results[best]}
        tmp = CACHE + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    return best, results
