"""Core + halo tiling (tiling.py:16-61, run.py:114-129 of the reference): windows, .npy tile
inputs and case tile jobs, on the CPU."""
import json

import numpy as np
import pytest

from solweig_amd import tiling


@pytest.mark.parametrize('width,height,core,halo', [(100, 70, 32, 10), (64, 64, 64, 5), (50, 41, 7, 0),
                                                   (33, 90, 40, 100)])
def test_cores_cover_the_raster_once_and_halos_clip(width, height, core, halo):
    tiles = tiling.tile_windows(width, height, core, halo)
    cover = np.zeros((height, width), np.int32)
    for t in tiles:
        col, row, w, h = t['core_window']
        cover[row:row + h, col:col + w] += 1
        left, top, iw, ih = t['input_window']
        assert 0 <= left and 0 <= top and left + iw <= width and top + ih <= height
        assert left == max(0, col - halo) and top == max(0, row - halo)
        assert left + iw == min(width, col + w + halo) and top + ih == min(height, row + h + halo)
        r0, c0, oh, ow = t['output_core']
        assert (r0 + top, c0 + left, oh, ow) == (row, col, h, w)
    assert (cover == 1).all()
    assert tiles[0]['tile_id'] == 'r000_c000'


def _write_rasters(folder, shape, landcover=False, seed=0):
    rng = np.random.default_rng(seed)
    arrays = {'building_dsm_path': rng.uniform(0, 30, shape).astype(np.float32),
              'dem_path': rng.uniform(0, 2, shape).astype(np.float32),
              'tree_path': rng.uniform(0, 10, shape).astype(np.float32)}
    if landcover:
        arrays['landcover_path'] = rng.integers(1, 7, shape).astype(np.float32)
    folder.mkdir(parents=True, exist_ok=True)
    for key, a in arrays.items():
        np.save(folder / tiling.INPUTS[key], a)
    return arrays


@pytest.mark.parametrize('fmt', ['npy', 'view'])
def test_create_tiles_windows_of_every_raster(tmp_path, fmt):
    src = _write_rasters(tmp_path / 'in', (45, 61), landcover=True)
    tiles = tiling.create_tiles(tmp_path / 'in', tmp_path / 'tiles', core=20, halo=6, format=fmt)
    assert len(tiles) == 3 * 4
    for t in tiles:
        assert set(t['inputs']) == set(src)
        left, top, w, h = t['input_window']
        for key, a in src.items():
            got = t['inputs'][key]
            got = np.load(got) if fmt == 'npy' else got
            np.testing.assert_array_equal(got, a[top:top + h, left:left + w])


def test_create_tiles_without_landcover_and_bad_shapes(tmp_path):
    src = _write_rasters(tmp_path / 'in', (30, 30))
    tiles = tiling.create_tiles(tmp_path / 'in', tmp_path / 'tiles', core=16, halo=4)
    assert all('landcover_path' not in t['inputs'] for t in tiles) and len(tiles) == 4
    np.save(tmp_path / 'in' / 'DEM.npy', src['dem_path'][:-1])
    with pytest.raises(ValueError):
        tiling.create_tiles(tmp_path / 'in', tmp_path / 'tiles', core=16, halo=4)
    with pytest.raises(ValueError):
        tiling.create_tiles(tmp_path / 'in', tmp_path / 'tiles', core=16, halo=4, format='vrt')


def test_case_tile_jobs(tmp_path):
    from solweig_amd import series
    _write_rasters(tmp_path / 'morph', (50, 50))
    case = {'building_dsm_path': 'morph/Building_DSM.npy', 'tree_path': 'morph/Trees.npy',
            'dem_path': 'morph/DEM.npy', 'met_file': 'met.txt', 'longitude': 0.0, 'latitude': 50.0,
            'core_area_m': 40, 'buffer_area_m': 10, 'resolution_m': 2}
    (tmp_path / 'case.json').write_text(json.dumps(case))
    c, settings = series.load_case(tmp_path / 'case.json')
    jobs = tiling.tile_jobs(c, settings, tmp_path / 'tiles')
    expect = tiling.tile_windows(50, 50, 20, 5)                 # 40 m / 2 m cores, 10 m / 2 m halos
    assert [label for label, _ in jobs] == [t['tile_id'] for t in expect]
    for (label, options), t in zip(jobs, expect):
        assert options['output_core'] == t['output_core']
        assert np.load(options['building_dsm_path']).shape == (t['input_window'][3], t['input_window'][2])
        assert options['met_file'] == c['met_file'] and options['latitude'] == 50.0
    single = tiling.tile_jobs(c, {'case_name': 'x'}, tmp_path / 'tiles')
    assert single == [('x', c)]
