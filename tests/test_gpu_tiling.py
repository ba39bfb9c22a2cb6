"""Tiled runs on the GPU: tiles whose halo covers the whole raster reproduce the untiled run on
their cores bit for bit, and a group of tiles run as one batch equals the tiles run one by one."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


def _inputs(tmp_path, d):
    from solweig_amd import tiling
    folder = tmp_path / 'morph'
    folder.mkdir()
    for key, name in (('building_dsm_path', 'dsm'), ('tree_path', 'trees'), ('dem_path', 'dem')):
        np.save(folder / tiling.INPUTS[key], getattr(d, name))
    return folder


def test_full_halo_tiles_match_the_untiled_run(rt, tmp_path):
    from solweig_amd import series, synthetic, tiling
    d = synthetic.make_c1(seed=31, size=72, days=1)
    folder = _inputs(tmp_path, d)
    # cores of 36 with a halo of 72: every tile's input window is the whole raster
    tiles = tiling.create_tiles(folder, tmp_path / 'tiles', core=36, halo=72)
    print(f'full-halo tiles: {len(tiles)}', flush=True)
    res = tiling.run_tiles(tiles, d.met, tmp_path / 'out', d.longitude, d.latitude, utc_offset_hours=0.0,
                           no_write=True)
    ref = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path / 'ref', d.longitude, d.latitude,
                                utc_offset_hours=0.0, no_write=True, return_arrays=True)
    (date, want), = ref['days'].items()
    for t in tiles:
        col, row, w, h = t['core_window']
        got = res[t['tile_id']]['days'][date]
        for name in ('TMRT', 'UTCI'):
            assert got[name].shape == (24, h, w)
            assert np.array_equal(got[name], want[name][:, row:row + h, col:col + w], equal_nan=True), (t, name)


def test_tile_batch_equals_tiles_one_by_one(rt, tmp_path):
    from solweig_amd import series, synthetic
    d = synthetic.make_c1(seed=32, size=64, days=1)
    folder = _inputs(tmp_path, d)
    np.savetxt(tmp_path / 'met.txt', d.met, header='met', comments='')
    case = {'building_dsm_path': str(folder / 'Building_DSM.npy'), 'tree_path': str(folder / 'Trees.npy'),
            'dem_path': str(folder / 'DEM.npy'), 'met_file': str(tmp_path / 'met.txt'),
            'longitude': d.longitude, 'latitude': d.latitude, 'utc_offset_hours': 0.0,
            'core_area_m': 40, 'buffer_area_m': 12}
    (tmp_path / 'case.json').write_text(json.dumps(case))
    print('tiled case: batched tiles', flush=True)
    batch = series.run_case(tmp_path / 'case.json', output=tmp_path / 'a', no_write=True)
    print('tiled case: one run per tile', flush=True)
    single = series.run_case(tmp_path / 'case.json', output=tmp_path / 'b', no_write=True, batch_tiles=False,
                             return_arrays=True)
    assert sorted(batch) == sorted(single) == ['r000_c000', 'r000_c001', 'r001_c000', 'r001_c001']
    for tile_id in batch:
        (date, a), = batch[tile_id]['days'].items()
        b = single[tile_id]['days'][date]
        core = 40 if tile_id.endswith('c000') else 24
        assert a['UTCI'].shape[2] == core
        for name in ('TMRT', 'UTCI'):
            assert np.array_equal(a[name], b[name], equal_nan=True), (tile_id, name)
        assert np.isfinite(a['TMRT']).all()
