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
