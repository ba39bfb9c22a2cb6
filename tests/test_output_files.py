"""Output files of the series driver (CPU): chunked day stacks, their band tags (the GeoTIFF tags
of output.py:132-157 as a .json beside each .npy), output window and valid mask."""
import json

import numpy as np

from solweig_amd import series


def _rows(hours, minutes=0, doy=172):
    m = np.full((len(hours), 24), -999.0)
    m[:, 0], m[:, 1], m[:, 2], m[:, 3] = 2025, doy, hours, minutes
    return m


def test_band_tags_follow_the_reference_contract():
    t = series.band_tags('TMRT', '2025-06-21', [0, 13], [0, 30], 1.0)
    assert t['variable'] == 'TMRT' and t['units'] == 'degree_Celsius' and t['time_basis'] == 'forcing_local_time'
    assert t['bands'][1] == {'description': 'TMRT_13:30_local', 'Time': '2025-06-21T13:30:00',
                             'time_basis': 'forcing_local_time', 'utc_offset_hours': '1.0',
                             'solar_sampling': 'forcing_time_minus_half_timestep', 'units': 'degree_Celsius'}
    assert series.band_tags('Kup', '2025-06-21', [5], [0], 0.0)['units'] == 'W m-2'
    assert series.band_tags('Shadow', '2025-06-21', [5], [0], 0.0)['units'] == '1'


def test_write_day_chunks_crops_and_masks(tmp_path):
    hours = list(range(24))
    rows = _rows(hours)
    rng = np.random.default_rng(0)
    stacks = {'UTCI': rng.random((24, 6, 7)).astype(np.float32), 'Kdown': rng.random((24, 6, 7)).astype(np.float32)}
    crop = (slice(1, 5), slice(2, 7))
    valid = np.ones((4, 5), bool)
    valid[0, 0] = False
    files = series.write_day(tmp_path / 'd', '2025-06-21', stacks, rows, 10, 0.0, crop=crop, valid=valid)
    assert [p.rsplit('/', 1)[1] for p in files['UTCI']] == ['UTCI_0001.npy', 'UTCI_0002.npy', 'UTCI_0003.npy']
    parts = [np.load(p) for p in files['Kdown']]
    assert [a.shape for a in parts] == [(10, 4, 5), (10, 4, 5), (4, 4, 5)]
    back = np.concatenate(parts)
    assert np.isnan(back[:, 0, 0]).all()
    ref = stacks['Kdown'][:, 1:5, 2:7].copy()
    ref[:, 0, 0] = np.nan
    np.testing.assert_array_equal(back, ref)
    tags = json.loads((tmp_path / 'd' / 'Kdown_0003.json').read_text())
    assert [b['description'] for b in tags['bands']] == [f'Kdown_{h:02d}:00_local' for h in range(20, 24)]
    # a day within one chunk keeps the plain name
    one = series.write_day(tmp_path / 'e', '2025-06-21', {'UTCI': stacks['UTCI']}, rows, 24, 0.0)
    assert one['UTCI'][0].endswith('/UTCI.npy') and np.load(one['UTCI'][0]).shape == (24, 6, 7)
