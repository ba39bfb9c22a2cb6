"""Core + halo tiling of aligned rasters (drop-in for tiling.create_tiles, tiling.py:16-61, and the
tile jobs of run.py:114-129), with a tile runner that puts many tiles into one launch.

A raster larger than one run is cut into cores that cover it once; each core is simulated with a
halo of surrounding pixels (clipped at the raster edge) and its outputs are cropped back to the
core (``output_core``). Rasters here are ``.npy`` files (the reference writes GDAL VRT or GeoTIFF
windows, which need GDAL): ``format="npy"`` writes each tile's input windows as ``.npy`` copies,
``format="view"`` keeps them as read-only memory-mapped windows of the sources.

``run_tiles`` is the MI355X form of the reference's tile loop: instead of one run per tile, the
tiles are grouped up to a pixel budget and every group runs as one batch of uneven domains of the
batched engine (one launch per pass and hour for the whole group), each tile's outputs written
cropped to its core as ``run_timeseries`` writes them.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

INPUTS = {
    'building_dsm_path': 'Building_DSM.npy',
    'dem_path': 'DEM.npy',
    'tree_path': 'Trees.npy',
    'landcover_path': 'Landcover.npy',
}


def tile_windows(width, height, core=1800, halo=100):
    """The tile list of create_tiles (tiling.py:28-43): row-major over cores of ``core`` pixels,
    ``core_window`` [col, row, width, height], ``input_window`` [left, top, width, height] (the
    core grown by ``halo`` and clipped to the raster), ``output_core`` [row, col, height, width]
    of the core inside the input window."""
    if core <= 0 or halo < 0:
        raise ValueError('core must be positive and halo non-negative')
    tiles = []
    for tile_row, row in enumerate(range(0, height, core)):
        for tile_col, col in enumerate(range(0, width, core)):
            core_width, core_height = min(core, width - col), min(core, height - row)
            left, top = max(0, col - halo), max(0, row - halo)
            right, bottom = min(width, col + core_width + halo), min(height, row + core_height + halo)
            tiles.append({
                'tile_id': f'r{tile_row:03d}_c{tile_col:03d}',
                'core_window': [col, row, core_width, core_height],
                'input_window': [left, top, right - left, bottom - top],
                'output_core': [row - top, col - left, core_height, core_width],
            })
    return tiles


def _window(array, win):
    left, top, w, h = win
    return array[top:top + h, left:left + w]


def create_tiles(input_dir, output_dir, *, core=1800, halo=100, format='npy'):
    """Split the aligned rasters of ``input_dir`` (Building_DSM, DEM, Trees and optionally
    Landcover, as .npy) into core + halo tiles (tiling.py:16-61). Returns the tile list of
    ``tile_windows`` with ``inputs``: per raster key the tile's window (a path under
    ``output_dir/<tile_id>/`` for ``format="npy"``, a read-only array for ``format="view"``)."""
    if format not in ('npy', 'view'):
        raise ValueError("format must be 'npy' or 'view' (GDAL formats need GDAL)")
    input_dir, output_dir = Path(input_dir).resolve(), Path(output_dir).resolve()
    sources = {}
    for key, filename in INPUTS.items():
        path = input_dir / filename
        if key == 'landcover_path' and not path.exists():
            continue
        sources[key] = np.load(path, mmap_mode='r', allow_pickle=False)
    height, width = sources['building_dsm_path'].shape
    for key, a in sources.items():
        if a.shape != (height, width):
            raise ValueError(f'{INPUTS[key]}: shape {a.shape} differs from the DSM {(height, width)}')
    tiles = tile_windows(width, height, core, halo)
    if format == 'npy':
        output_dir.mkdir(parents=True, exist_ok=True)
    for tile in tiles:
        inputs = {}
        for key, a in sources.items():
            win = _window(a, tile['input_window'])
            if format == 'view':
                inputs[key] = win
                continue
            folder = output_dir / tile['tile_id']
            folder.mkdir(exist_ok=True)
            target = folder / INPUTS[key]
            np.save(target, np.ascontiguousarray(win))
            inputs[key] = str(target)
        tile['inputs'] = inputs
    return tiles


def tile_jobs(case, settings, tiles_dir):
    """(label, run_timeseries keyword arguments) per tile of a case (run.py:114-129): one job with
    the whole case when ``core_area_m`` is not set; otherwise core and halo are ``core_area_m`` and
    ``buffer_area_m`` over the pixel size (``pixel_size_m``, else ``resolution_m``)."""
    if 'core_area_m' not in settings:
        return [(settings.get('case_name', 'tile'), dict(case))]
    res = float(case.get('pixel_size_m', settings.get('resolution_m', 1)))
    core, halo = (int(round(settings[k] / res)) for k in ('core_area_m', 'buffer_area_m'))
    source_dir = Path(case['building_dsm_path']).parent
    tiles = create_tiles(source_dir, tiles_dir, core=core, halo=halo)
    keys = ['building_dsm_path', 'tree_path', 'dem_path'] + (['landcover_path'] if case.get('landcover_path') else [])
    return [(t['tile_id'], {**case, **{k: t['inputs'][k] for k in keys}, 'output_core': t['output_core']})
            for t in tiles]


# Device bytes that one pixel of a prepared domain keeps for the whole run (static.PreparedDomain):
# rasters and maps 37 (dsm, trees, walls, aspect, albedo, emissivity, TgK, Tstart, TmaxLST f32,
# ground u8), visibility bits 60, sky-view derived planes 24 (svfbuveg, svfveg f32, tan(asvf),
# tan(svfalfa) f64), neighbour record, mask byte and lag state 33, static sky sums 40, march canopy
# plane 4, ground-view planes 80 (step counts, unshadowed albedo sums, in-raster step counts).
STATIC_BYTES_PER_PIXEL = 278
# Default pixel budget of one batch: at 24 h and 2 output slots about 134 GB of device memory and
# 31 GB of page-locked host memory (batch_bytes).
DEFAULT_BATCH_PIXELS = 80_000_000


def batch_bytes(pixels, n_hours=24, slots=2):
    """(device bytes, page-locked host bytes) that a batch of ``pixels`` input pixels holds while
    it runs: the static state, the shadow pass's per-hour records and mask bytes and the diffuse
    sums (17 B per hour, in the two sets of the pipelined days), the day's Tmrt and UTCI (4 B each
    per hour) and, per output slot, one device staging copy and one page-locked host copy of them
    (output.DayOutputs). The sky-view pass adds 128 B per pixel of the domain being prepared while
    it runs."""
    per_day = 8 * int(n_hours)
    device = int(pixels) * (STATIC_BYTES_PER_PIXEL + 2 * 17 * int(n_hours) + per_day * (1 + int(slots)))
    host = int(pixels) * per_day * int(slots)
    return device, host


def _groups(tiles, max_pixels):
    group, size = [], 0
    for t in tiles:
        p = t['input_window'][2] * t['input_window'][3]
        if group and size + p > max_pixels:
            yield group
            group, size = [], 0
        group.append(t)
        size += p
    if group:
        yield group


def run_tiles(tiles, met_file, output_path, longitude, latitude, *, utc_offset_hours=None, site_altitude_m=3.,
              dsm_wall_threshold_m=3., pixel_size_m=1.0, save_tmrt=True, no_write=False, output_queue_size=2,
              max_batch_pixels=DEFAULT_BATCH_PIXELS, output_chunk_steps=24, progress=False):
    """Simulate every tile of ``tiles`` (``create_tiles`` output) over the forcing, the tiles of a
    group (up to ``max_batch_pixels`` input pixels; ``batch_bytes`` gives the device and
    page-locked host memory a group holds) as one batch of the batched engine. Outputs per
    tile and day, cropped to the tile's core: ``output_path/<tile_id>/<date>/UTCI.npy`` (and
    TMRT.npy), as run_timeseries writes them with ``output_core``. Returns {tile_id: {'days':
    {date: {var: [path]}}}} (with ``no_write``, the cropped host arrays instead of paths)."""
    from datetime import datetime
    from . import batched
    from .output import DayOutputs
    from .series import _raster, load_meteorology, make_domain, write_day
    from .sun import met_sun_device
    forcing = load_meteorology(met_file)
    utc = float(longitude) / 15.0 if utc_offset_hours is None else float(utc_offset_hours)
    out = Path(output_path).resolve()
    variables = ['UTCI'] + (['TMRT'] if save_tmrt else [])
    results = {t['tile_id']: {'days': {}} for t in tiles}

    rows_of = {d.date: d.values for d in forcing.days}

    def writer_for(tile):
        r0, c0, h, w = tile['output_core']
        window = (slice(r0, r0 + h), slice(c0, c0 + w))

        def write(date, tmrt, utci):
            stacks = {name: stack for name, stack in (('UTCI', utci), ('TMRT', tmrt)) if name in variables}
            if no_write:
                day = {name: np.array(stack[(slice(None),) + window]) for name, stack in stacks.items()}
            else:
                day = write_day(out / tile['tile_id'] / date, date, stacks, rows_of[date], output_chunk_steps, utc,
                                crop=window)
            results[tile['tile_id']]['days'][date] = day
        return write

    for group in _groups(tiles, max_batch_pixels):
        doms = [make_domain(_raster(t['inputs']['building_dsm_path']), _raster(t['inputs']['tree_path']),
                            _raster(t['inputs']['dem_path']), forcing, latitude, longitude, utc, site_altitude_m,
                            pixel_size_m, dsm_wall_threshold_m,
                            landcover=_raster(t['inputs'].get('landcover_path'))) for t in group]
        writes = [writer_for(t) for t in group]
        engine, outs = None, []
        try:
            days = forcing.days

            def tables_of(engine, day):
                sun = met_sun_device(day.values, doms[0].location, utc)
                return engine.tables(met=[day.values] * len(group), sun=[sun] * len(group))

            for i, day in enumerate(days):
                n = day.values.shape[0]
                if progress:
                    print(f'[{datetime.now():%H:%M:%S}] {day.date}: {len(group)} tiles, {n} steps', flush=True)
                if engine is None or engine.n_hours != n:
                    for o in outs:
                        o.close()
                    engine = batched.Engine(doms, n_hours=n, pipelined=True)
                    outs = [DayOutputs(engine, k, slots=max(1, int(output_queue_size))) for k in range(len(group))]
                # consecutive days of one engine are pipelined: the next day's shadow pass runs
                # beside this day's pixel pass (Engine.day_start / day_next)
                if not engine.day_open:
                    engine.day_start(tables_of(engine, day))
                nxt = days[i + 1] if i + 1 < len(days) and days[i + 1].values.shape[0] == n else None
                engine.day_next(None if nxt is None else tables_of(engine, nxt))
                for o, w in zip(outs, writes):
                    o.submit(day.date, w)
        finally:
            for o in outs:
                o.close()
    for r in results.values():
        r['days'] = dict(sorted(r['days'].items()))
    return results
