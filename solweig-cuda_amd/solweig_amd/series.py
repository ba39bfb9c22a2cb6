"""Time-series driver (drop-in for series.run_timeseries / load_meteorology, series.py:91-194,
and run.load_case, run.py:70-86).

Rasters are read from .npy files or passed as arrays (GeoTIFF needs GDAL, which this build does
not use); walls and aspect are computed from the DSM on the device (walls.calculate_dsm_walls,
series.py:212-236) and rounded to float32 as the reference stores them. The day loop runs on the
batched engine: one upload of the day's tables and 24 x 2 launches per day, the lag state reset
at each day start, UTCI NaN outside the ground mask (series.py:466-469). Outputs leave through
output.DayOutputs: staged on the device, copied to page-locked memory on a side stream and written
by a worker thread while the next days run. With ``plain=True`` every hour goes through the
per-step API (model.Solweig_2022a_calc) instead.
"""
from __future__ import annotations

from dataclasses import dataclass
from datetime import datetime, timedelta
import json
import os
from pathlib import Path

import numpy as np

from . import batched, synthetic
from . import walls as dsm_walls
from .output import DayOutputs
from .sun import met_sun


@dataclass(frozen=True)
class DayForcing:
    date: str
    values: np.ndarray
    timestamps: tuple


@dataclass(frozen=True)
class Meteorology:
    values: np.ndarray
    timestamps: tuple
    days: tuple
    timestep_minutes: float


def _read_values(source):
    if isinstance(source, np.ndarray):
        values = np.array(source, dtype=np.float64, order='C', copy=True)
        return values.reshape(1, -1) if values.ndim == 1 else values
    paths = [Path(source)] if isinstance(source, (str, os.PathLike)) else [Path(p) for p in source]
    arrays = [np.loadtxt(p.expanduser().resolve(), skiprows=1, ndmin=2, dtype=np.float64, encoding='utf-8-sig')
              for p in paths]
    return np.ascontiguousarray(arrays[0] if len(arrays) == 1 else np.concatenate(arrays), dtype=np.float64)


def load_meteorology(source) -> Meteorology:
    """Forcing table grouped by calendar date (series.py:175-194)."""
    if isinstance(source, Meteorology):
        return source
    values = _read_values(source)
    stamps = [datetime(int(y), 1, 1) + timedelta(days=int(d) - 1, hours=int(h), minutes=int(m))
              for y, d, h, m in values[:, :4].astype(np.int64)]
    timestep = (stamps[1] - stamps[0]).total_seconds() / 60 if len(stamps) > 1 else 0.0
    values.setflags(write=False)
    groups, start = [], 0
    for stop in range(1, len(stamps) + 1):
        if stop == len(stamps) or stamps[stop].date() != stamps[start].date():
            groups.append(DayForcing(stamps[start].date().isoformat(), values[start:stop], tuple(stamps[start:stop])))
            start = stop
    return Meteorology(values, tuple(stamps), tuple(groups), timestep)


def _raster(src):
    if src is None:
        return None
    if isinstance(src, np.ndarray):
        return src
    p = str(src)
    if p.endswith('.npy'):
        return np.load(p, allow_pickle=False)
    raise ValueError(f'{p}: only .npy rasters are read (GeoTIFF input needs GDAL)')


def make_domain(dsm, trees, dem, forcing, latitude, longitude, utc, site_altitude_m=3., pixel_size_m=1.0,
                dsm_wall_threshold_m=3., landcover=None):
    """The model domain of one raster set (series.py:212-365): float32 rasters, walls and aspect
    from the DSM on the device (rounded to float32 as series.py:235-236 stores them)."""
    dsm = np.asarray(dsm, dtype=np.float32)
    trees = np.asarray(trees, dtype=np.float32)
    dem = np.asarray(dem, dtype=np.float32)
    walls, aspect = dsm_walls.calculate_dsm_walls(dsm.astype(np.float64), float(pixel_size_m),
                                                  float(dsm_wall_threshold_m))
    walls, aspect = walls.astype(np.float32), aspect.astype(np.float32)
    return synthetic.Domain(dsm, trees, dem, walls, aspect, forcing.values, float(latitude), float(longitude),
                            float(utc), float(site_altitude_m), scale=1.0 / float(pixel_size_m),
                            landcover=None if landcover is None else np.asarray(landcover))


def run_timeseries(building_dsm_path, tree_path, dem_path, met_file, output_path, longitude, latitude, *,
                   landcover_path=None, utc_offset_hours=None, site_altitude_m=3., wall_method='dsm',
                   dsm_wall_threshold_m=3., save_tmrt=True, dtype=np.float32, no_write=False,
                   output_core=None, output_chunk_steps=24, async_output=True, output_queue_size=2,
                   output_valid_mask=None, progress=False, pixel_size_m=1.0, plain=False, **unused):
    """Run every day of the forcing; returns {'static': {...}, 'days': {date: {var: [paths]}}}
    (series.py:91-156). Outputs per day: UTCI (and TMRT) stacks [hours, rows, cols] as .npy."""
    if wall_method != 'dsm':
        raise ValueError("only wall_method='dsm' is supported")
    forcing = load_meteorology(met_file)
    utc = float(longitude) / 15.0 if utc_offset_hours is None else float(utc_offset_hours)
    dom = make_domain(_raster(building_dsm_path), _raster(tree_path), _raster(dem_path), forcing, latitude,
                      longitude, utc, site_altitude_m, pixel_size_m, dsm_wall_threshold_m,
                      landcover=_raster(landcover_path))
    out = Path(output_path).resolve()
    files = {'static': {}, 'days': {}}
    variables = ['UTCI'] + (['TMRT'] if save_tmrt else [])
    crop = (slice(None), slice(None)) if output_core is None else \
        (slice(output_core[0], output_core[0] + output_core[2]), slice(output_core[1], output_core[1] + output_core[3]))
    valid = None if output_valid_mask is None else np.asarray(output_valid_mask, dtype=bool)

    def write(date, tmrt, utci):
        if valid is not None:
            utci = np.where(valid[None], utci, np.nan)
        folder = out / date
        folder.mkdir(parents=True, exist_ok=True)
        day_files = {v: [] for v in variables}
        for name, stack in (('UTCI', utci), ('TMRT', tmrt)):
            if name in variables:
                p = folder / f'{name}.npy'
                np.save(p, stack[(slice(None),) + crop])
                day_files[name].append(str(p))
        files['days'][date] = day_files

    engine = writer = None
    try:
        for day in forcing.days:
            n = day.values.shape[0]
            if progress:
                print(f'[{datetime.now():%H:%M:%S}] {day.date}: {n} steps', flush=True)
            sun = met_sun(day.values, dom.location, utc)
            if plain:
                tmrt, utci = _plain_day(dom, day.values, sun)
            else:
                if engine is None or engine.n_hours != n:
                    if writer is not None:
                        writer.close()
                    engine = batched.Engine([dom], n_hours=n)
                    writer = None if no_write else DayOutputs(engine, slots=max(1, int(output_queue_size)))
                engine.load(engine.tables(met=[day.values], sun=[sun]))
                engine.reset_state()
                engine.forward_day()
                if writer is not None:
                    writer.submit(day.date, write)
                    continue
                tmrt, utci = engine.outputs(0)
            if no_write:
                if valid is not None:
                    utci = np.where(valid[None], utci, np.nan)
                files['days'][day.date] = {'UTCI': utci[(slice(None),) + crop], 'TMRT': tmrt[(slice(None),) + crop]}
            else:
                write(day.date, tmrt, utci)
    finally:
        if writer is not None:
            writer.close()
    files['days'] = dict(sorted(files['days'].items()))
    return files


def _plain_day(dom, met, sun):
    """One day through the per-step path (the plain form, series.py:445-469)."""
    from . import model
    from .static import PreparedDomain
    pd = PreparedDomain(dom)
    prep = model.prepare_from_domain(pd, dom.location, dom.scale)
    return plain_day(prep, met, sun, ground=pd.dev['ground'])


def plain_day(prep, met, sun, ground=None, keep=True):
    """Hour loop of series._run_day over model.Solweig_2022a_calc-style steps; returns host
    (Tmrt, UTCI) stacks (``keep``) or leaves the last hour on the device."""
    from . import model, utci as u
    n = met.shape[0]
    dectime = sun['dectime']
    timestepdec = (dectime[1] - dectime[0]) if n > 1 else 1 / 24.
    CI, fdt, timeadd = 1., 1, 0.
    lags = [0.0] * 6
    tm, ut = [], []
    for i in range(n):
        if dectime[i] - np.floor(dectime[i]) == 0:
            CI = 1.
        meta, (CI, fdt, timeadd, _) = model.run_step(
            prep, altitude=sun['altitude'][i], azimuth=sun['azimuth'][i], zen=sun['zen'][i], jday=sun['jday'][i],
            dectime=dectime[i], altmax=sun['altmax'][i], Ta=met[i, 11], RH=met[i, 10], radG=met[i, 14],
            P=met[i, 12], psi=0.03 if 97 < met[i, 1] < 300 else 0.5, CI=CI, firstdaytime=fdt, timeadd=timeadd,
            timestepdec=timestepdec, lags=lags, Ws=met[i, 9])
        lags = [prep.domain.dev['lag'].plane(k) for k in range(6)]
        tmrt = prep.tmrt.plane(0, prep.shape)
        utci = u.utci_calculator(met[i, 11], met[i, 10], tmrt, met[i, 9], ground=ground, stream=prep.stream)
        prep.stream.sync()
        if keep:
            tm.append(tmrt.to_host())
            ut.append(utci.to_host())
    return (np.stack(tm), np.stack(ut)) if keep else (tmrt, utci)


PATH_KEYS = {'building_dsm_path', 'tree_path', 'dem_path', 'met_file', 'building_geometry_path',
             'building_height_path', 'landcover_path', 'output_path'}
PREPARATION_KEYS = {'case_name', 'case_type', 'case_size_m', 'core_area_m', 'buffer_area_m', 'resolution_m',
                    'start_date', 'end_date'}


def load_case(path):
    """case.json -> (run_timeseries kwargs, preparation settings) (run.py:70-86)."""
    path = Path(path).resolve()
    case = json.loads(path.read_text(encoding='utf-8-sig'))
    settings = {k: case.pop(k) for k in PREPARATION_KEYS if k in case}
    settings.setdefault('case_type', 'local')
    settings.setdefault('resolution_m', 1)
    for key in PATH_KEYS & case.keys():
        value = case[key]
        if value is None:
            continue
        entries = value if isinstance(value, list) else [value]
        resolved = [str((path.parent / Path(e).expanduser()).resolve()) for e in entries]
        case[key] = resolved if isinstance(value, list) else resolved[0]
    case.setdefault('async_output', True)
    case.setdefault('output_queue_size', 2)
    case.setdefault('output_chunk_steps', 24)
    return case, settings


def run_case(case_path, *, output=None, no_write=False, progress=False, batch_tiles=True, **overrides):
    """run.py main for a local case (run.py:20-67): the case.json fields (``overrides`` replace
    them), and with ``core_area_m`` / ``buffer_area_m`` the spatial tiles of run.py:114-129, each
    written under ``output/<tile_id>``. ``batch_tiles`` runs the tiles as batches of one launch
    (tiling.run_tiles) instead of one run per tile. Returns the files dict of run_timeseries, or
    {tile_id: files} for a tiled case."""
    from . import tiling
    case, settings = load_case(case_path)
    case.update(overrides)
    if settings.get('case_type', 'local') != 'local':
        raise ValueError('only local cases are run here (open-data acquisition needs the network)')
    default = Path(case_path).resolve().parent / 'output'
    output = Path(output or case.pop('output_path', default)).resolve()
    case.pop('output_path', None)
    forcing = load_meteorology(case['met_file'])
    tiles_dir = Path(case_path).resolve().parent / 'tiles'
    if 'core_area_m' in settings and batch_tiles:
        res = float(case.get('pixel_size_m', settings.get('resolution_m', 1)))
        core, halo = (int(round(settings[k] / res)) for k in ('core_area_m', 'buffer_area_m'))
        tiles = tiling.create_tiles(Path(case['building_dsm_path']).parent, tiles_dir, core=core, halo=halo)
        keep = {k: case[k] for k in ('utc_offset_hours', 'site_altitude_m', 'dsm_wall_threshold_m', 'save_tmrt',
                                     'output_queue_size') if k in case}
        return tiling.run_tiles(tiles, forcing, output, case['longitude'], case['latitude'], pixel_size_m=res,
                                no_write=no_write, progress=progress, **keep)
    results = {}
    for label, options in tiling.tile_jobs(case, settings, tiles_dir):
        destination = output / label if 'core_area_m' in settings else output
        results[label] = run_timeseries(**{**options, 'met_file': forcing}, output_path=destination,
                                        no_write=no_write, progress=progress)
    return results if 'core_area_m' in settings else results[next(iter(results))]
