"""Time-series driver (drop-in for series.run_timeseries / load_meteorology, series.py:91-194,
and run.load_case, run.py:70-86).

Rasters are read from .npy files or passed as arrays (GeoTIFF needs GDAL, which is not used
here); walls and aspect are computed from the DSM on the device (walls.calculate_dsm_walls,
series.py:212-236) and rounded to float32 as the reference stores them. The day loop runs on the
batched engine: one upload of the day's tables and the day's launches, the lag state reset
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
from .sun import met_sun_device


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


UNITS = {'UTCI': 'degree_Celsius', 'TMRT': 'degree_Celsius', 'Shadow': '1', 'Kup': 'W m-2', 'Kdown': 'W m-2',
         'Lup': 'W m-2', 'Ldown': 'W m-2'}
MODEL_TAG = 'SOLWEIG 2022a (solweig_amd, MI355X)'


def band_tags(variable, date, hours, minutes, utc):
    """The dataset and per-band tags of one output stack, as the reference writes them into its
    GeoTIFF (output.py:132-157): band description ``<variable>_HH:MM_local``, ISO time of the
    forcing row, time basis, UTC offset, solar-sampling rule and units."""
    units = UNITS[variable]
    base = datetime.fromisoformat(date)
    bands = []
    for h, m in zip(hours, minutes):
        h, m = int(h), int(m)
        bands.append({'description': f'{variable}_{h:02d}:{m:02d}_local',
                      'Time': base.replace(hour=h, minute=m).isoformat(), 'time_basis': 'forcing_local_time',
                      'utc_offset_hours': str(utc), 'solar_sampling': 'forcing_time_minus_half_timestep',
                      'units': units})
    return {'model': MODEL_TAG, 'variable': variable, 'units': units, 'time_basis': 'forcing_local_time',
            'bands': bands}


def write_day(folder, date, stacks, met_rows, chunk_steps, utc, crop=None, valid=None):
    """Write each variable's [hours, rows, cols] stack of one day under ``folder``: files of at most
    ``chunk_steps`` bands, suffixed _0001, _0002, ... when the day has more rows than that
    (series.py:472-489), each ``<name>.npy`` beside a ``<name>.json`` of its tags (band_tags).
    ``crop`` (row slice, col slice) is the output window; ``valid`` a mask of the window's shape,
    NaN written where it is False in every variable (output.py:205-213). Returns {var: [paths]}."""
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    n = met_rows.shape[0]
    c = max(1, int(chunk_steps))
    crop = (slice(None), slice(None)) if crop is None else tuple(crop)
    files = {}
    for var, stack in stacks.items():
        paths = []
        for k, lo in enumerate(range(0, n, c)):
            hi = min(n, lo + c)
            values = np.array(stack[lo:hi][(slice(None),) + crop], dtype=np.float32)
            if valid is not None:
                values[:, ~valid] = np.nan
            name = f'{var}_{k + 1:04d}' if n > c else var
            np.save(folder / f'{name}.npy', values)
            tags = band_tags(var, date, met_rows[lo:hi, 2], met_rows[lo:hi, 3], utc)
            (folder / f'{name}.json').write_text(json.dumps(tags, indent=1), encoding='utf-8')
            paths.append(str(folder / f'{name}.npy'))
        files[var] = paths
    return files


def _write_static(folder, name, values, variable, units, extra=None):
    folder.mkdir(parents=True, exist_ok=True)
    np.save(folder / f'{name}.npy', np.asarray(values, dtype=np.float32))
    tags = {'model': MODEL_TAG, 'variable': variable, 'units': units, **(extra or {})}
    (folder / f'{name}.json').write_text(json.dumps(tags, indent=1), encoding='utf-8')
    return str(folder / f'{name}.npy')


def run_timeseries(building_dsm_path, tree_path, dem_path, met_file, output_path, longitude, latitude, *,
                   landcover_path=None, utc_offset_hours=None, site_altitude_m=3., wall_method='dsm',
                   dsm_wall_threshold_m=3., save_tmrt=True, save_svf=False, save_kup=False, save_kdown=False,
                   save_lup=False, save_ldown=False, save_shadow=False, dtype=np.float32, no_write=False,
                   output_core=None, output_chunk_steps=24, async_output=True, output_queue_size=2,
                   output_valid_mask=None, progress=False, pixel_size_m=1.0, plain=False, return_arrays=False,
                   table_workers=None, **unused):
    """Run every day of the forcing; returns {'static': {...}, 'days': {date: {var: [paths]}}}
    (series.py:91-156). Outputs per day: UTCI, and TMRT, Kup, Kdown, Lup, Ldown, Shadow as
    enabled, as [hours, rows, cols] .npy stacks split into ``output_chunk_steps`` bands with a .json
    of their tags; static: walls and aspect, and SVF (svfbuveg, series.py:343-378) with
    ``save_svf``. ``no_write`` writes nothing and leaves the days empty, as the reference does;
    ``return_arrays`` (with ``no_write``) returns the host stacks instead. Forcing of more than two
    days builds the day tables ahead on ``table_workers`` CPU processes (default: host cores - 1,
    at most 8; 0 builds them in this process). ``dtype=np.float64`` (run.py:28) runs every step
    through the reference's decomposition in FP64 (stepwise_day); outputs are float32 files in
    either precision (output.py:119)."""
    if wall_method != 'dsm':
        raise ValueError("only wall_method='dsm' is supported")
    dtype = np.dtype(dtype)
    if dtype not in (np.float32, np.float64):
        raise ValueError('dtype must be float32 or float64')
    fp64 = dtype == np.float64
    forcing = load_meteorology(met_file)
    utc = float(longitude) / 15.0 if utc_offset_hours is None else float(utc_offset_hours)
    dom = make_domain(_raster(building_dsm_path), _raster(tree_path), _raster(dem_path), forcing, latitude,
                      longitude, utc, site_altitude_m, pixel_size_m, dsm_wall_threshold_m,
                      landcover=_raster(landcover_path))
    out = Path(output_path).resolve()
    files = {'static': {}, 'days': {}}
    extras = [name for name, on in (('Kup', save_kup), ('Kdown', save_kdown), ('Lup', save_lup),
                                    ('Ldown', save_ldown), ('Shadow', save_shadow)) if on]
    variables = ['UTCI'] + (['TMRT'] if save_tmrt else []) + extras
    crop = (slice(None), slice(None)) if output_core is None else \
        (slice(output_core[0], output_core[0] + output_core[2]), slice(output_core[1], output_core[1] + output_core[3]))
    valid = None if output_valid_mask is None else np.asarray(output_valid_mask, dtype=bool)
    if not no_write:
        static = out / 'static'
        files['static']['walls'] = _write_static(static, 'walls', dom.walls, 'wall_height', 'm',
                                                 {'wall_method': 'dsm', 'wall_backend': 'HIP'})
        files['static']['aspect'] = _write_static(static, 'aspect', dom.aspect, 'wall_aspect',
                                                  'degree_clockwise_from_north', {'wall_method': 'dsm'})
    rows_of = {d.date: d.values for d in forcing.days}

    def stacks_of(tmrt, utci, extra=None):
        got = {'UTCI': utci, 'TMRT': tmrt, **(extra or {})}
        return {v: got[v] for v in variables}

    def write(date, tmrt, utci, extra=None):
        files['days'][date] = write_day(out / date, date, stacks_of(tmrt, utci, extra), rows_of[date],
                                        output_chunk_steps, utc, crop, valid)

    days = forcing.days
    prefetch = len(days) > 2 and not plain and not fp64 and table_workers != 0
    engine = writer = steps = None
    pending = {}
    try:
        for i, day in enumerate(days):
            n = day.values.shape[0]
            if progress:
                print(f'[{datetime.now():%H:%M:%S}] {day.date}: {n} steps', flush=True)
            extra = None
            if fp64:
                steps = steps or stepwise_context(dom)
                sun = met_sun_device(day.values, dom.location, utc)
                tmrt, utci, extra = stepwise_day(*steps, day.values, sun, extras)
            elif plain:
                sun = met_sun_device(day.values, dom.location, utc)
                tmrt, utci, extra = _plain_day(dom, day.values, sun, extras)
            else:
                if engine is None or engine.n_hours != n:
                    if writer is not None:
                        writer.close()
                    if engine is not None:
                        engine.close_table_workers()
                    engine = batched.Engine([dom], n_hours=n, extra_outputs=bool(extras),
                                            keep_outputs=not no_write or return_arrays, pipelined=True)
                    pending = {}
                    workers = engine.start_table_workers(table_workers) if prefetch else 0
                    if save_svf and not no_write and 'SVF' not in files['static']:
                        pd = engine.domains[0]
                        svfbuveg = pd._svf - (1.0 - pd._svfveg) * (1.0 - pd.trans)
                        files['static']['SVF'] = _write_static(out / 'static', 'SVF', svfbuveg, 'SVF', '1')
                    writer = None if no_write else DayOutputs(engine, slots=max(1, int(output_queue_size)))
                # the tables of the next days are built on the workers while this one runs
                for j in range(i, min(len(days), i + 2 + (workers if prefetch else 0))):
                    if j not in pending and days[j].values.shape[0] == n:
                        pending[j] = engine.tables_async(met=[days[j].values])
                # consecutive days of one engine are pipelined: the next day's shadow pass runs
                # beside this day's pixel pass (Engine.day_start / day_next)
                if not engine.day_open:
                    engine.day_start(pending.pop(i).result())
                nxt = pending.pop(i + 1).result() if i + 1 in pending else None
                engine.day_next(nxt)
                if writer is not None:
                    writer.submit(day.date, write)
                    continue
                if not return_arrays:
                    files['days'][day.date] = {}
                    continue
                tmrt, utci = engine.outputs(0)
                extra = engine.extra_outputs(0) if extras else None
            if no_write and not return_arrays:
                files['days'][day.date] = {}
            elif no_write:
                day_stacks = {}
                for v, st in stacks_of(tmrt, utci, extra).items():
                    a = np.array(st[(slice(None),) + crop], dtype=np.float32)
                    if valid is not None:
                        a[:, ~valid] = np.nan
                    day_stacks[v] = a
                files['days'][day.date] = day_stacks
            else:
                write(day.date, tmrt, utci, extra)
    finally:
        if writer is not None:
            writer.close()
        if engine is not None:
            engine.stream.sync()
            engine.close_table_workers()
    files['days'] = dict(sorted(files['days'].items()))
    return files


def _plain_day(dom, met, sun, extras=()):
    """One day through the per-step path (the plain form, series.py:445-469): (Tmrt, UTCI,
    {extra variable: stack} or None)."""
    from . import model
    from .static import PreparedDomain
    pd = PreparedDomain(dom)
    prep = model.prepare_from_domain(pd, dom.location, dom.scale)
    return plain_day(prep, met, sun, ground=pd.dev['ground'], extras=extras)


def plain_day(prep, met, sun, ground=None, keep=True, extras=None):
    """Hour loop of series._run_day over model.Solweig_2022a_calc-style steps; returns host
    (Tmrt, UTCI) stacks (``keep``) or leaves the last hour on the device. With ``extras`` (a list
    of rt.EXTRA_NAMES) the third value holds their host stacks from the full-mode planes."""
    from . import model, utci as u
    n = met.shape[0]
    dectime = sun['dectime']
    timestepdec = (dectime[1] - dectime[0]) if n > 1 else 1 / 24.
    CI, fdt, timeadd = 1., 1, 0.
    lags = [0.0] * 6
    tm, ut, ex = [], [], {}
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
            for name in extras or ():
                ex.setdefault(name, []).append(prep.full.plane(FULL_OF[name], prep.shape).to_host())
    if extras is not None:
        return np.stack(tm), np.stack(ut), ({k: np.stack(v) for k, v in ex.items()} if extras else None)
    return (np.stack(tm), np.stack(ut)) if keep else (tmrt, utci)


STEP_OF = {'Kdown': 'Kdown', 'Kup': 'Kup', 'Ldown': 'Ldown', 'Lup': 'Lup', 'Shadow': 'shadow'}
STEP_CONSTANTS = ('albedo_b', 'absK', 'absL', 'ewall', 'Fside', 'Fup', 'Fcyl', 'cyl', 'elvis', 'onlyglobal',
                  'TgK_wall', 'Tstart_wall', 'TmaxLST_wall')


def stepwise_context(dom, dtype=np.float64):
    """(PreparedDomain, static inputs, stepwise.PreparedSteps) of a domain for stepwise_day: the
    static preparation of series._prepare_spatial (series.py:243-369) with the model prepared in
    ``dtype``."""
    from . import model
    from .static import PreparedDomain, model_inputs
    pd = PreparedDomain(dom, keep_host=True)
    static = model_inputs(pd)
    return pd, static, model.prepare_model(dtype=dtype, stepwise=True, **static)


def stepwise_day(pd, static, prep, met, sun, extras=()):
    """One day through the reference's step decomposition (stepwise.step, the dtype of ``prep``),
    the hour loop of series._run_day (series.py:396-469): host float32 stacks (Tmrt, UTCI,
    {extra variable: stack} or None). UTCI takes the float32-rounded Tmrt and is NaN off the ground
    mask, as utci_calculator and series.py:468-469 make it."""
    from . import stepwise, utci as u
    n = met.shape[0]
    dectime = sun['dectime']
    timestepdec = (dectime[1] - dectime[0]) if n > 1 else 1 / 24.
    CI, fdt, timeadd = 1., 1., 0.
    state = dict.fromkeys(('Tgmap1', 'Tgmap1E', 'Tgmap1S', 'Tgmap1W', 'Tgmap1N', 'TgOut1'), 0.0)
    twater = float(np.mean(met[:, 11])) if static['landcover'] else None
    consts = {k: static[k] for k in STEP_CONSTANTS}
    tm, ut, ex = [], [], {name: [] for name in extras}
    for i in range(n):
        if dectime[i] - np.floor(dectime[i]) == 0:
            CI = 1.
        o = stepwise.step(prep, altitude=sun['altitude'][i], azimuth=sun['azimuth'][i], zen=sun['zen'][i],
                          jday=sun['jday'][i], dectime=dectime[i], altmax=sun['altmax'][i], Ta=met[i, 11],
                          RH=met[i, 10], radG=met[i, 14], radD=met[i, 21], radI=met[i, 22], P=met[i, 12],
                          psi=0.03 if 97 < met[i, 1] < 300 else 0.5, CI=CI, firstdaytime=fdt, timeadd=timeadd,
                          timestepdec=timestepdec, location=static['location'], landcover=static['landcover'],
                          Twater=twater, **state, **consts)
        CI, fdt, timeadd = o['CI'], o['firstdaytime'], o['timeadd']
        state = {k: o[k] for k in state}
        t32 = stepwise.as_float32(o['Tmrt'], prep)
        utci = u.utci_calculator(met[i, 11], met[i, 10], t32, met[i, 9], ground=pd.dev['ground'])
        tm.append(t32.to_host())
        ut.append(utci.to_host())
        for name in extras:
            ex[name].append(stepwise.as_float32(o[STEP_OF[name]], prep).to_host())
    return np.stack(tm), np.stack(ut), ({k: np.stack(v) for k, v in ex.items()} if extras else None)


FULL_OF = {'Kdown': 0, 'Kup': 1, 'Ldown': 2, 'Lup': 3, 'Shadow': 19}   # rt.FULL_PLANES positions

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


def run_case(case_path, *, output=None, no_write=False, progress=False, batch_tiles=True, return_arrays=False,
             **overrides):
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
                                        no_write=no_write, progress=progress, return_arrays=return_arrays)
    return results if 'core_area_m' in settings else results[next(iter(results))]
