"""Per-step drop-in of the reference model API (model.py:15-256, 961-1016).

``prepare_model`` uploads the static fields once; ``Solweig_2022a_calc`` runs one SOLWEIG 2022a
step with the reference's signature and returns its 39-tuple in ``OUTPUT_NAMES`` order. A step
is the plain per-step form: the host computes the hour's scalars and tables (sky.hour_tables),
uploads them, and two device launches write every per-pixel output plane (the shadow pass in
full mode and the pixel pass in full mode); the lag maps are read from and written back to the
caller's state arrays. Per-pixel planes come back as device views (``.get()`` copies to host).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import runtime as rt
from . import sky
from .static import prepared_from_arrays

SBC = 5.67051e-8

OUTPUT_NAMES = ('Tmrt', 'Kdown', 'Kup', 'Ldown', 'Lup', 'Tg', 'ea', 'esky', 'I0', 'CI',
                'shadow', 'firstdaytime', 'timestepdec', 'timeadd', 'Tgmap1', 'Tgmap1E', 'Tgmap1S',
                'Tgmap1W', 'Tgmap1N', 'Keast', 'Ksouth', 'Kwest', 'Knorth', 'Least', 'Lsouth', 'Lwest',
                'Lnorth', 'KsideI', 'TgOut1', 'TgOut', 'radI', 'radD', 'Lside', 'L_patches', 'CI_Tg',
                'CI_TgG', 'KsideD', 'dRad', 'Kside')
STATE_NAMES = ('CI', 'firstdaytime', 'timeadd', 'Tgmap1', 'Tgmap1E', 'Tgmap1S', 'Tgmap1W', 'Tgmap1N', 'TgOut1')
FIELD_NAMES = ('svf', 'svfN', 'svfW', 'svfE', 'svfS', 'svfveg', 'svfNveg', 'svfEveg', 'svfSveg', 'svfWveg',
               'svfaveg', 'svfEaveg', 'svfSaveg', 'svfWaveg', 'svfNaveg', 'lc_grid', 'TgK', 'Tstart',
               'alb_grid', 'emis_grid', 'TmaxLST', 'svfbuveg', 'asvf', 'svfalfa')
LAG_NAMES = ('Tgmap1', 'Tgmap1E', 'Tgmap1S', 'Tgmap1W', 'Tgmap1N', 'TgOut1')


@dataclass
class PreparedModel:
    domain: object
    hours: object
    ptab_f: object
    ptab_d: object
    ray_off: object
    ray_dz: object
    full: object
    shadow_planes: object
    tmrt: object
    dom_dev: object
    block_start: object
    gvf_off: object
    gvf_bounds: object
    first: int
    second: int
    shape: tuple
    dtype: np.dtype
    stream: object
    location: dict
    scale: float


def _host(v):
    if isinstance(v, (rt.DeviceArray, rt.DeviceView)):
        return v.to_host()
    return np.asarray(v.get() if hasattr(v, 'get') else v)


def prepare_model(*, dtype=np.float32, stepwise=False, **inputs):
    """Upload the static fields of series._prepare_spatial's dict (model.py:987-1016). FP32 (the
    default) prepares the fused hour passes; ``dtype=np.float64``, or ``stepwise=True`` in FP32,
    prepares the reference's decomposition of the step into its sub-model calls
    (stepwise.PreparedSteps), which Solweig_2022a_calc then runs."""
    dtype = np.dtype(dtype)
    if dtype == np.float64 or stepwise:
        from .stepwise import prepare_stepwise
        return prepare_stepwise(dtype=dtype, **inputs)
    if dtype != np.float32:
        raise ValueError('dtype must be float32 or float64')
    if not int(inputs.get('anisotropic_sky', 1)) or int(inputs.get('patch_option', 2)) != 2 \
            or not int(inputs.get('usevegdem', 1)) or not int(inputs.get('cyl', 1)):
        raise ValueError('supported configuration: cyl=1, anisotropic_sky=1, usevegdem=1, patch_option=2')
    rows, cols = int(inputs['rows']), int(inputs['cols'])
    g = {k: _host(inputs[k]) for k in ('dsm', 'vegdem', 'walls', 'dirwalls', 'buildings', 'svfbuveg', 'asvf',
                                       'svfalfa', 'shmat', 'vegshmat', 'vbshvegshmat', 'alb_grid', 'emis_grid',
                                       'TgK', 'Tstart', 'TmaxLST')}
    stream = rt.Stream()
    dom = prepared_from_arrays(
        dsm=g['dsm'], vegdem=g['vegdem'], walls=g['walls'], dirwalls=g['dirwalls'], buildings=g['buildings'],
        svfbuveg=g['svfbuveg'], asvf=g['asvf'], svfalfa=g['svfalfa'], shmat=g['shmat'], vegshmat=g['vegshmat'],
        vbshvegshmat=g['vbshvegshmat'], alb_grid=g['alb_grid'], emis_grid=g['emis_grid'], TgK=g['TgK'],
        Tstart=g['Tstart'], TmaxLST=g['TmaxLST'], landcover=int(inputs.get('landcover', 0)),
        wall_params=(float(inputs['TgK_wall']), float(inputs['Tstart_wall']), float(inputs['TmaxLST_wall'])),
        amaxvalue=float(inputs['amaxvalue']), scale=float(inputs['scale']), stream=stream)
    loc = inputs.get('location') or {'latitude': 0.0}
    return prepare_from_domain(dom, loc, float(inputs['scale']), float(inputs.get('first', 1.0)),
                               float(inputs.get('second', 22.0)), stream=stream)


def prepare_from_domain(dom, location, scale=1.0, first=1.0, second=22.0, stream=None):
    """Per-step buffers around an already prepared domain (static.PreparedDomain)."""
    stream = stream or rt.Stream()
    rows, cols = dom.shape
    P = rows * cols
    cap = max(rows, cols) + 2 + sky.STEP_BLOCK
    bufs = dict(hours=rt.DeviceArray((1, rt.SW_HOUR_STRIDE), np.float64),
                ptab_f=rt.DeviceArray((1, rt.SW_PATCHES, rt.SW_PATCH_F), np.float32),
                ptab_d=rt.DeviceArray((1, rt.SW_PATCH_D, rt.SW_PATCHES), np.float64),
                ray_off=rt.DeviceArray((cap, 2), np.int32), ray_dz=rt.DeviceArray((cap,), np.float64),
                full=rt.DeviceArray((len(rt.FULL_PLANES), P), np.float32),
                shadow_planes=rt.DeviceArray((8, P), np.float64), tmrt=rt.DeviceArray((1, P), np.float32))
    s = dom.struct(hours=bufs['hours'].ptr, ptab_f=bufs['ptab_f'].ptr, ptab_d=bufs['ptab_d'].ptr,
                   ray_off=bufs['ray_off'].ptr, ray_dz=bufs['ray_dz'].ptr, tmrt=bufs['tmrt'].ptr,
                   full=bufs['full'].ptr, shadow_planes=bufs['shadow_planes'].ptr)
    dom_dev = rt.DeviceArray((ctypes.sizeof(s),), np.uint8)
    rt.call('sw_memcpy_h2d', dom_dev.ptr, ctypes.addressof(s), ctypes.sizeof(s), stream.handle)
    stream.sync()
    bs = rt.DeviceArray.from_host(np.array([0, dom.tiles], dtype=np.int32))
    off, bounds, f_steps, s_steps = sky.gvf_geometry(scale, first, second)
    return PreparedModel(dom, bufs['hours'], bufs['ptab_f'], bufs['ptab_d'], bufs['ray_off'], bufs['ray_dz'],
                         bufs['full'], bufs['shadow_planes'], bufs['tmrt'], dom_dev, bs,
                         rt.DeviceArray.from_host(off, np.int32), rt.DeviceArray.from_host(bounds, np.float64),
                         f_steps, s_steps, (rows, cols), np.dtype(np.float32), stream, location, float(scale))


def _set_lag(prep, k, value):
    P = prep.shape[0] * prep.shape[1]
    lag = prep.domain.dev['lag']
    dst = lag.ptr + k * P * 4
    if isinstance(value, (rt.DeviceArray, rt.DeviceView)) and value.dtype == np.float32:
        if value.ptr != dst:
            rt.call('sw_memcpy_d2d', dst, value.ptr, P * 4, prep.stream.handle)
        return
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(_host(value), dtype=np.float32), prep.shape))
    rt.call('sw_memcpy_h2d', dst, a.ctypes.data, a.nbytes, prep.stream.handle)
    prep.stream.sync()


def _lag_copy(prep, k):
    P = prep.shape[0] * prep.shape[1]
    out = rt.DeviceArray(prep.shape, np.float32)
    rt.call('sw_memcpy_d2d', out.ptr, prep.domain.dev['lag'].ptr + k * P * 4, P * 4, prep.stream.handle)
    return out


def run_step(prep, *, altitude, azimuth, zen, jday, dectime, altmax, Ta, RH, radG, P, psi, CI,
             firstdaytime, timeadd, timestepdec, lags, Ws=-999.0, shadow_planes=True):
    """Host tables + the two full-mode launches of one step. Returns (meta, new state)."""
    dom = prep.domain
    h, pf, pd, off, dz, st, meta = sky.hour_tables(
        float(altitude), float(azimuth), float(zen), float(jday), float(dectime), float(altmax), float(Ta),
        float(RH), float(radG), float(P), float(Ws), float(jday), float(prep.location['latitude']),
        prep.shape, prep.scale, dom.amaxvalue, (float(CI), int(firstdaytime), float(timeadd), float(timestepdec)),
        wall_params=dom.wall_params, psi=float(psi))
    h[rt.H['RAY_START']] = 0
    h[rt.H['PATCH_TABLE']] = 0
    prep.hours.upload(h[None], prep.stream)
    prep.ptab_f.upload(pf[None], prep.stream)
    prep.ptab_d.upload(np.ascontiguousarray(pd.T)[None], prep.stream)
    if len(dz):
        off, dz = sky.pad_steps(off, dz)
        prep.ray_off.upload(off, prep.stream)
        prep.ray_dz.upload(dz, prep.stream)
    for k, v in enumerate(lags):
        _set_lag(prep, k, v)
    rt.call('sw_hour_shadow', prep.dom_dev.ptr, prep.block_start.ptr, 1, dom.tiles, 0, int(shadow_planes),
            prep.stream.handle)
    rt.call('sw_hour_pixels', prep.dom_dev.ptr, prep.block_start.ptr, 1, dom.tiles, 0, prep.gvf_off.ptr,
            prep.gvf_bounds.ptr, prep.first, prep.second, 1, prep.stream.handle)
    return meta, st


def Solweig_2022a_calc(i, dsm, scale, rows, cols, svf, svfN, svfW, svfE, svfS, svfveg, svfNveg, svfEveg, svfSveg,
                       svfWveg, svfaveg, svfEaveg, svfSaveg, svfWaveg, svfNaveg, vegdem, vegdem2, albedo_b, absK,
                       absL, ewall, Fside, Fup, Fcyl, altitude, azimuth, zen, jday, usevegdem, onlyglobal,
                       buildings, location, psi, landcover, lc_grid, dectime, altmax, dirwalls, walls, cyl,
                       elvis, Ta, RH, radG, radD, radI, P, amaxvalue, bush, Twater, TgK, Tstart, alb_grid,
                       emis_grid, TgK_wall, Tstart_wall, TmaxLST, TmaxLST_wall, first, second, svfalfa,
                       svfbuveg, firstdaytime, timeadd, timestepdec, Tgmap1, Tgmap1E, Tgmap1S, Tgmap1W,
                       Tgmap1N, CI, TgOut1, diffsh, shmat, vegshmat, vbshvegshmat, anisotropic_sky, asvf,
                       patch_option, *, dtype=np.float32, prepared=None):
    """One SOLWEIG 2022a step (model.py:15-256): returns the 39-tuple of OUTPUT_NAMES. With a
    stepwise preparation (dtype float64, or prepare_model(stepwise=True)) the step runs as the
    reference's sequence of sub-model calls in that precision; otherwise as the two fused FP32
    passes."""
    from .stepwise import PreparedSteps, step as stepwise_step
    if prepared is None and np.dtype(dtype) == np.float64:
        args = {k: v for k, v in locals().items() if k not in ('prepared', 'PreparedSteps', 'stepwise_step')}
        prepared = prepare_model(**args)
    if isinstance(prepared, PreparedSteps):
        out = stepwise_step(prepared, altitude=altitude, azimuth=azimuth, zen=zen, jday=jday, dectime=dectime,
                            altmax=altmax, Ta=Ta, RH=RH, radG=radG, radD=radD, radI=radI, P=P, psi=psi, CI=CI,
                            firstdaytime=firstdaytime, timeadd=timeadd, timestepdec=timestepdec, Tgmap1=Tgmap1,
                            Tgmap1E=Tgmap1E, Tgmap1S=Tgmap1S, Tgmap1W=Tgmap1W, Tgmap1N=Tgmap1N, TgOut1=TgOut1,
                            location=location, landcover=int(float(landcover)), Twater=Twater,
                            albedo_b=float(albedo_b), absK=float(absK), absL=float(absL), ewall=float(ewall),
                            Fside=float(Fside), Fup=float(Fup), Fcyl=float(Fcyl), cyl=int(cyl), elvis=float(elvis),
                            onlyglobal=int(onlyglobal), TgK_wall=float(TgK_wall), Tstart_wall=float(Tstart_wall),
                            TmaxLST_wall=float(TmaxLST_wall))
        return tuple(out[n] for n in OUTPUT_NAMES)
    consts = (float(albedo_b), float(ewall), float(absK), float(absL), float(Fside), float(Fup), float(Fcyl))
    if consts != (0.2, 0.9, 0.7, 0.95, 0.22, 0.06, 0.28) or float(elvis) != 0 or int(onlyglobal) != 1:
        raise ValueError('the device path is built for the series.py constants (albedo_b 0.2, ewall 0.9, '
                         'absK 0.7, absL 0.95, Fside 0.22, Fup 0.06, Fcyl 0.28, elvis 0, onlyglobal 1)')
    if prepared is None:
        args = {k: v for k, v in locals().items() if k not in ('prepared', 'consts', 'PreparedSteps',
                                                               'stepwise_step')}
        prepared = prepare_model(**args)
    prep = prepared
    prep.location = location
    meta, (CI2, fdt2, timeadd2, _) = run_step(
        prep, altitude=altitude, azimuth=azimuth, zen=zen, jday=jday, dectime=dectime, altmax=altmax, Ta=Ta,
        RH=RH, radG=radG, P=P, psi=psi, CI=CI, firstdaytime=firstdaytime, timeadd=timeadd,
        timestepdec=timestepdec, lags=(Tgmap1, Tgmap1E, Tgmap1S, Tgmap1W, Tgmap1N, TgOut1))
    F = {name: prep.full.plane(k, prep.shape) for k, name in enumerate(rt.FULL_PLANES)}
    lags = [_lag_copy(prep, k) for k in range(6)]
    Tmrt = prep.tmrt.plane(0, prep.shape)
    day = float(altitude) > 0
    L_patches = None
    if day:
        alt, azi, lv = sky.perez(float(zen) * (180 / np.pi), float(azimuth), meta['radD'], meta['radI'], float(jday))
        L_patches = np.column_stack((alt, azi, lv))
    else:
        a, z = sky.patches()
        L_patches = np.column_stack((a, z, np.zeros_like(a)))
    out = dict(Tmrt=Tmrt, ea=meta['ea'], esky=meta['esky'], I0=meta['I0'], CI=CI2, firstdaytime=fdt2,
               timestepdec=timestepdec, timeadd=timeadd2, TgOut1=lags[5], radI=meta['radI'] if day else radI,
               radD=meta['radD'] if day else radD, L_patches=L_patches, CI_Tg=meta['CI_Tg'],
               CI_TgG=meta['CI_TgG'])
    for k, name in enumerate(LAG_NAMES[:5]):
        out[name] = lags[k]
    for name in ('Kdown', 'Kup', 'Ldown', 'Lup', 'Tg', 'shadow', 'Keast', 'Ksouth', 'Kwest', 'Knorth', 'Least',
                 'Lsouth', 'Lwest', 'Lnorth', 'KsideI', 'TgOut', 'Lside', 'KsideD', 'dRad', 'Kside'):
        out[name] = F[name]
    return tuple(out[n] for n in OUTPUT_NAMES)


@dataclass
class PreparedShadow:
    """Device state of the sun-ray march of one raster (model.py:626-652)."""
    dev: dict
    shape: tuple
    building_max: float
    vegetation_max: float
    scale: float
    usevegdem: bool
    hours: object
    ray_off: object
    ray_dz: object
    planes: object
    dom_dev: object
    block_start: object
    tiles: int
    stream: object


def prepare_shadow(a, walls, aspect, scale, *, vegdem=None, vegdem2=None, bush=None, amaxvalue=None):
    """Static state of cast_shadow (model.py:638-652). ``aspect`` in radians as the step passes it
    (float32 degrees x pi/180 in FP64); ``a`` and the canopy top ``vegdem`` FP32-exact;
    ``vegdem2``/``bush`` must be the ones series.py:282-287 derives from them."""
    from .static import height_bound
    a = np.asarray(_host(a), dtype=np.float64)
    rows, cols = a.shape
    P = rows * cols
    if P > rt.SW_MAX_PIXELS:
        raise ValueError(f'prepare_shadow: a {rows}x{cols} raster exceeds the {rt.SW_MAX_PIXELS}-pixel limit')
    dsm32 = a.astype(np.float32)
    if not np.array_equal(dsm32.astype(np.float64), a):
        raise ValueError('prepare_shadow: the DSM must be FP32-exact')
    if not (a.min() >= 0.0):
        # the march skips windows against a wave-wide lower bound of non-negative thresholds and
        # treats off-raster steps as no-ops, both exact only for heights >= 0
        raise ValueError('prepare_shadow: the DSM has heights below 0; raise it to a minimum of 0 first, '
                         'as the series preparation does (series.py:252-257)')
    use = vegdem is not None
    if use:
        top = np.asarray(_host(vegdem), dtype=np.float64)
        trees = np.where(top != 0, top - a, 0.0)
        trees32 = trees.astype(np.float32)
        if np.any(trees32 < 0) or not np.array_equal(np.where(top != 0, trees32.astype(np.float64) + a, 0.0), top):
            raise ValueError('prepare_shadow: vegdem must be an FP32-exact canopy top above the DSM')
        v2 = trees32.astype(np.float64) * 0.25 + a
        v2[v2 == a] = 0
        if vegdem2 is not None and not np.array_equal(np.asarray(_host(vegdem2), dtype=np.float64), v2):
            raise ValueError('prepare_shadow: vegdem2 differs from 0.25 trees + dsm (series.py:285-286)')
        if bush is not None and not np.array_equal(np.asarray(_host(bush), dtype=np.float64),
                                                   np.logical_not(v2 * top) * top):
            raise ValueError('prepare_shadow: bush differs from not(vegdem2 vegdem) vegdem (series.py:287)')
    else:
        trees32 = np.zeros((rows, cols), np.float32)
    asp = np.asarray(_host(aspect), dtype=np.float64)
    deg32 = (asp * (180.0 / np.pi)).astype(np.float32)
    if not np.array_equal(deg32.astype(np.float64) * np.pi / 180.0, asp):
        raise ValueError('prepare_shadow: aspect must be float32 degrees times pi/180 (model.py:995-996)')
    stream = rt.Stream()
    up = lambda x, t=np.float32: rt.DeviceArray.from_host(np.ascontiguousarray(np.broadcast_to(x, (rows, cols)), dtype=t))
    zero = up(np.float32(0))
    dev = dict(dsm=up(dsm32), trees=up(trees32), walls=up(np.asarray(_host(walls), np.float32)), aspect=up(deg32),
               ground=up(np.uint8(1), np.uint8), zero=zero, one=up(np.float32(1)),
               tan64=rt.DeviceArray((P,), np.float64), rec=rt.DeviceArray((P, 2), np.float32),
               hmax=rt.DeviceArray.from_host(height_bound(a, trees32)))
    building_max = float(a.max())
    veg_max = float(amaxvalue) if (use and amaxvalue is not None) else float(max(a.max() - a.min(), trees32.max()))
    cap = max(rows, cols) + 2 + sky.STEP_BLOCK
    hours = rt.DeviceArray((1, rt.SW_HOUR_STRIDE), np.float64)
    ray_off, ray_dz = rt.DeviceArray((cap, 2), np.int32), rt.DeviceArray((cap,), np.float64)
    planes = rt.DeviceArray((8, rows, cols), np.float64)
    s = rt.SwDomain()
    s.rows, s.cols, s.pixels, s.landcover, s.tile_cols = rows, cols, P, 0, rt.TILE_X
    for f in ('emis', 'alb', 'tgk', 'tstart', 'svfbuveg'):
        setattr(s, f, dev['zero'].ptr)
    s.tmaxlst = dev['one'].ptr
    s.dsm, s.trees, s.walls, s.aspect = dev['dsm'].ptr, dev['trees'].ptr, dev['walls'].ptr, dev['aspect'].ptr
    s.ground, s.rec, s.hmax, s.hmax_cols = dev['ground'].ptr, dev['rec'].ptr, dev['hmax'].ptr, -(-cols // rt.HMAX_CELL)
    s.tan_asvf = s.tan_svfalfa = dev['tan64'].ptr
    s.hours, s.ray_off, s.ray_dz, s.shadow_planes = hours.ptr, ray_off.ptr, ray_dz.ptr, planes.ptr
    dom_dev = rt.DeviceArray((ctypes.sizeof(s),), np.uint8)
    rt.call('sw_memcpy_h2d', dom_dev.ptr, ctypes.addressof(s), ctypes.sizeof(s), stream.handle)
    tiles = rt.tile_count((rows, cols))
    bs = rt.DeviceArray.from_host(np.array([0, tiles], dtype=np.int32))
    stream.sync()
    return PreparedShadow(dev, (rows, cols), building_max, veg_max, float(scale), use, hours, ray_off, ray_dz,
                          planes, dom_dev, bs, tiles, stream)


def cast_shadow(geometry, azimuth, altitude):
    """Shadows of one sun position (model.py:689-708): with vegetation the 8 FP64 planes (vegsh,
    sh, vbshvegsh, wallsh, wallsun, wallshve, facesh, facesun), else (sh, wallsh, wallsun, facesh,
    facesun), as device views. Masks are bit-identical to shadowingfunction_wallheight_23."""
    if isinstance(geometry, PreparedModel):
        raise TypeError('cast_shadow takes the state of prepare_shadow')
    g = geometry
    amax = g.vegetation_max if g.usevegdem else g.building_max
    off, dz, low, high, middle = sky.shadow_ray(float(azimuth), float(altitude), g.scale, amax, g.shape)
    h = np.zeros((1, rt.SW_HOUR_STRIDE))
    h[0, rt.H['DAY']] = 1.0
    h[0, rt.H['RAY_START']], h[0, rt.H['RAY_LEN']] = 0, len(dz)
    h[0, rt.H['FACE_LOW']], h[0, rt.H['FACE_HIGH']], h[0, rt.H['FACE_MIDDLE']] = low, high, float(middle)
    h[0, rt.H['PSI']] = 1.0
    g.hours.upload(h, g.stream)
    if len(dz):
        off, dz = sky.pad_steps(off, dz)
        g.ray_off.upload(off, g.stream)
        g.ray_dz.upload(dz, g.stream)
    rt.call('sw_hour_shadow', g.dom_dev.ptr, g.block_start.ptr, 1, g.tiles, 0, 1, g.stream.handle)
    g.stream.sync()
    planes = tuple(g.planes.plane(k, g.shape) for k in range(8))
    return planes if g.usevegdem else tuple(planes[k] for k in (1, 3, 4, 6, 7))


# per-step sub-models with the reference's names (model.py:736-958)
from .flux import (Kup_veg_2015a, Lside_veg_v2022a, Perez_v3, TsWaveDelay_2015a, assemble_tmrt,  # noqa: E402,F401
                   cylindric_wedge, surface_temperature)
from .groundview import GvfGeometry, GvfResult, gvf, prepare_gvf_geometry  # noqa: E402,F401
from .anisotropic import (Kside_veg_v2022a, PreparedGeometry, PreparedShortwave, longwave,  # noqa: E402,F401
                          prepare_geometry, prepare_shortwave)
