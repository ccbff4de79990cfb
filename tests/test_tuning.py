"""The tile-shape and batch-execution cache of hip/tuning.py (no device)."""
import types

from sar_pixel_simulator_amd.hip import tuning


def _model(nz, ny, nx):
    return types.SimpleNamespace(sim=types.SimpleNamespace(shape=(nz, ny, nx)))


def test_batch_key_is_order_independent():
    a = [_model(67, 101, 101), _model(133, 267, 267)]
    assert tuning.batch_key(a) == tuning.batch_key(a[::-1])
    assert tuning.batch_key(a) == 'batch:gfx950:101x101x67;267x267x133'


def test_batch_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(tuning, 'CACHE', str(tmp_path / 'cache.json'))
    models = [_model(35, 31, 33), _model(37, 39, 45)]
    assert tuning.batch_mode_for(models) is None
    data = {tuning.batch_key(models): {'label': 'serial-4', 'mode': [1, 0, 4], 'launch': None,
                                       'ms_per_step': 1.0},
            tuning.batch_key(models[:1]): {'label': 'packed-z4', 'mode': 'packed',
                                           'launch': [0, 4, 0, 0, 0], 'ms_per_step': 2.0}}
    import json
    with open(tuning.CACHE, 'w') as f:
        json.dump(data, f)
    assert tuning.batch_mode_for(models) == ((1, 0, 4), None)
    assert tuning.batch_mode_for(models[:1]) == ('packed', (0, 4, 0, 0, 0))


def test_default_stepper_by_grid_size(monkeypatch):
    from sar_pixel_simulator_amd.hip import common
    monkeypatch.setattr(common, 'FUSED_STEPPER', None)
    # the batched soil sizes and the serving canopy sizes stay on the split
    # stepper; the single large run and the largest domain take the halo one
    assert common.stepper_for(267, 267, 133) == 0
    assert common.stepper_for(381, 381, 311) == 0
    assert common.stepper_for(761, 761, 203) == 3
    assert common.stepper_for(1523, 1523, 311) == 3
    n = common.HALO_MIN_CELLS
    assert common.stepper_for(n, 1, 1) == 3 and common.stepper_for(n - 1, 1, 1) == 0
    monkeypatch.setattr(common, 'FUSED_STEPPER', 0)
    assert common.stepper_for(761, 761, 203) == 0


def test_every_stepper_has_a_default_shape():
    for fused, name in tuning.STEPPERS.items():
        assert name in tuning.DEFAULT and name in tuning.CANDIDATES
        assert len(tuning.launch_for(1, 1, 1, fused=fused)) == 5
    # the halo stepper's candidates are the split stepper's block shapes
    # with at most 8 wavefronts and 1 or 2 cells per lane
    for ty, zc, xv, xw, xo in tuning.CANDIDATES['halo']:
        assert xv in (1, 2) and ty * xw <= 8
        assert xo in (0, 3, 4, 5)           # plain order or fine bands of 1, 2, 4 row blocks
    # flat tiles are a split-stepper shape of one cell per lane
    assert all(xv == 1 and ty == 0 for ty, zc, xv, xw, xo in tuning.CANDIDATES['split'] if xw == -1)
