"""The per-step forward in the reference's own decomposition (model.py:15-256, 987-1016).

``prepare_stepwise`` keeps the static state of the step as the reference's ``prepare_model`` does:
the 24 FIELD_NAMES planes on the device (asvf, svfalfa and lc_grid in FP64, the rest in the
working precision), the shadow, ground-view, day and night patch geometries and the shortwave
diffuse transmissivities. ``step`` then runs one SOLWEIG 2022a step as a sequence of standalone
device calls in the order of Solweig_2022a_calc: host scalars (daylen, Prata, clearness, Reindl,
Perez), dRad, cast_shadow, the shadow composition, surface temperature, gvf, the five Lup wave
delays and TgOut's, cylindric_wedge, Kdown, Kup_veg_2015a, Kside_veg_v2022a, longwave,
Lside_veg_v2022a and assemble_tmrt, each a kernel of libsolweig_hip in FP32 or FP64 with the
caller's state planes threaded through. It is the plain per-step GPU forward of the reference's
structure (about 20 launches and host round trips per daytime hour), the ``dtype=float64`` form of
the drop-in API, and the device FP64 reference against which the fused FP32 hour passes are
audited at full raster sizes.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace

import numpy as np

from . import anisotropic, flux, groundview, model, sky
from . import runtime as rt
from .flux import field_table, working_dtype

SBC = 5.67051e-8
FP64_FIELDS = ('asvf', 'svfalfa', 'lc_grid')


@dataclass
class PreparedSteps:
    fields: dict
    shadow: object
    gvf: object
    longwave: object
    night_longwave: object
    shortwave: object
    sky: np.ndarray
    dtype: np.dtype
    shape: tuple
    usevegdem: int


def _host(v):
    if isinstance(v, (rt.DeviceArray, rt.DeviceView)):
        return v.to_host()
    return np.asarray(v.get() if hasattr(v, 'get') else v)


def _field(value, shape, dtype):
    a = np.asarray(_host(value), dtype=dtype)
    if a.size == 1:
        a = np.broadcast_to(a.reshape(()), shape)
    return rt.DeviceArray.from_host(np.ascontiguousarray(a.reshape(shape), dtype=dtype))


def prepare_stepwise(*, dtype=np.float32, **inputs):
    """The static state of the step (model.py:987-1016) for the inputs of series._prepare_spatial's
    dict (model.FIELD_NAMES, dsm, walls, dirwalls, buildings, vegdem, vegdem2, bush, amaxvalue,
    scale, first, second, shmat, vegshmat, vbshvegshmat, diffsh, usevegdem, patch_option)."""
    dtype = working_dtype(dtype)
    shape = (int(inputs['rows']), int(inputs['cols']))
    if int(inputs.get('anisotropic_sky', 1)) != 1 or int(inputs.get('patch_option', 2)) != 2:
        raise ValueError('the step is built for anisotropic_sky=1 and patch_option=2 (153 patches)')
    veg = int(inputs.get('usevegdem', 1))
    fields = {n: _field(inputs[n], shape, np.float64 if n in FP64_FIELDS else dtype) for n in model.FIELD_NAMES}
    dirwalls = np.asarray(_host(inputs['dirwalls']), dtype=np.float64)
    shadow = model.prepare_shadow(inputs['dsm'], inputs['walls'], dirwalls * np.pi / 180., inputs['scale'],
                                  vegdem=inputs['vegdem'] if veg else None,
                                  vegdem2=inputs['vegdem2'] if veg else None,
                                  bush=inputs['bush'] if veg else None,
                                  amaxvalue=inputs['amaxvalue'] if veg else None)
    ground = groundview.prepare_gvf_geometry(inputs['walls'], inputs['buildings'], dirwalls, inputs['scale'],
                                             inputs['first'], inputs['second'])
    alt, azi = sky.patches()
    sky_ = np.column_stack((alt, azi, np.zeros_like(alt)))
    day_sky = sky_.copy()
    day_sky[:, :2] = day_sky[:, :2] * (np.pi / 180.) * (180. / np.pi)      # model.py:1007-1008
    lw = anisotropic.prepare_geometry(day_sky, inputs['shmat'], inputs['vegshmat'], inputs['vbshvegshmat'])
    night = replace(lw, sky_patches=sky_, steradian=sky.steradians(alt),
                    directional_cosines=anisotropic.directional_cosines(azi))
    sw = anisotropic.prepare_shortwave(lw, inputs['diffsh'], dtype=dtype)
    return PreparedSteps(fields, shadow, ground, lw, night, sw, sky_, dtype, shape, veg)


def _terms(op, values, params, prep, fp64_fields=(), out_dtype=None):
    n = prep.shape[0] * prep.shape[1]
    D = prep.dtype if out_dtype is None else np.dtype(out_dtype)
    P, C, keep = field_table(values, n, D, fp64_fields)
    out = rt.DeviceArray(prep.shape, D)
    Q = (ctypes.c_double * max(1, len(params)))(*[float(x) for x in params])
    f = rt.lib().sw_step_terms
    vp, ci = ctypes.c_void_p, ctypes.c_int
    f.argtypes = [ci, vp, vp, ci, ctypes.c_longlong, vp, ci, ci, vp, vp]
    f.restype = ci
    rt.check(f(int(op), P, C, len(values), n, Q, len(params), D.itemsize, out.ptr, None))
    rt.check(rt.lib().sw_stream_sync(None))
    del keep
    return out


def _diffuse(prep, lv, radD):
    """dRad (model.py:1019-1027)."""
    n = prep.shape[0] * prep.shape[1]
    lum = rt.DeviceArray.from_host(np.ascontiguousarray(np.asarray(lv)[:, 2], dtype=np.float64))
    out = rt.DeviceArray(prep.shape, prep.dtype)
    f = rt.lib().sw_diffuse
    vp, ci = ctypes.c_void_p, ctypes.c_int
    f.argtypes = [vp, vp, ctypes.c_longlong, ci, ctypes.c_double, ci, vp, vp]
    f.restype = ci
    rt.check(f(prep.shortwave.diffuse_visibility.ptr, lum.ptr, n, len(lv), float(radD), prep.dtype.itemsize, out.ptr,
               None))
    rt.check(rt.lib().sw_stream_sync(None))
    return out


def _state(v, prep):
    """A lag-state plane in the working precision (a device array is passed through)."""
    if isinstance(v, (rt.DeviceArray, rt.DeviceView)) and v.dtype == prep.dtype and tuple(v.shape) == prep.shape:
        return v
    return _field(v, prep.shape, prep.dtype)


def step(prep, *, altitude, azimuth, zen, jday, dectime, altmax, Ta, RH, radG, radD, radI, P, psi, CI,
         firstdaytime, timeadd, timestepdec, Tgmap1, Tgmap1E, Tgmap1S, Tgmap1W, Tgmap1N, TgOut1, location,
         landcover=0, Twater=None, albedo_b=0.2, absK=0.7, absL=0.95, ewall=0.9, Fside=0.22, Fup=0.06, Fcyl=0.28,
         cyl=1, elvis=0.0, onlyglobal=1, TgK_wall=0.37, Tstart_wall=-3.41, TmaxLST_wall=15.0):
    """One SOLWEIG 2022a step (model.py:15-256) -> dict of OUTPUT_NAMES (device planes and host
    scalars)."""
    f = prep.fields
    D = prep.dtype
    opts = dict(dtype=D)
    rows, cols = prep.shape
    altitude, azimuth, zen, jday, Ta, RH, radG, radD, radI, P, CI, altmax, dectime = (
        np.float64(float(v)) for v in (altitude, azimuth, zen, jday, Ta, RH, radG, radD, radI, P, CI, altmax,
                                       dectime))
    t = 0.
    states = [_state(v, prep) for v in (Tgmap1, Tgmap1E, Tgmap1S, Tgmap1W, Tgmap1N, TgOut1)]
    Tgmap1, Tgmap1E, Tgmap1S, Tgmap1W, Tgmap1N, TgOut1 = states
    SNUP = sky.daylen(jday, location['latitude'])
    ea = 6.107 * 10 ** ((7.5 * Ta) / (237.3 + Ta)) * (RH / 100.)
    msteg = 46.5 * (ea / (Ta + 273.15))
    esky = 1 - (1 + msteg) * np.exp(-((1.2 + 3.0 * msteg) ** 0.5)) + elvis
    zero = None
    if altitude > 0:
        I0, CI, Kt = sky.clearness(zen, jday, Ta, RH / 100., radG, location['latitude'], P)
        if (CI > 1) or (CI == np.inf):
            CI = 1
        if onlyglobal == 1:
            radI, radD = sky.diffuse_fraction(radG, altitude, Kt, Ta, RH)
        lv, _, _ = flux.Perez_v3(zen * (180 / np.pi), azimuth, radD, radI, jday, 1, 2)
        dRad = _diffuse(prep, lv, radD)
        cast = model.cast_shadow(prep.shadow, azimuth, altitude)
        if prep.usevegdem:
            vegsh, sh, _, wallsh, wallsun, wallshve, _, facesun = cast
            shadow = _terms(0, (sh, vegsh), (psi,), prep, fp64_fields=(0, 1))
        else:
            sh, wallsh, wallsun, _, facesun = cast
            shadow = _terms(0, (sh, 1.0), (psi,), prep, fp64_fields=(0, 1))
        radI0, radD0 = sky.diffuse_fraction(I0, altitude, 1., Ta, RH)
        corr = 0.1473 * np.log(90 - zen / np.pi * 180) + 0.3454
        CI_Tg = radG / radI0 + (1 - corr)
        if (CI_Tg > 1) or (CI_Tg == np.inf):
            CI_Tg = 1
        radG0 = radI0 * np.sin(altitude * (np.pi / 180)) + radD0
        CI_TgG = radG / radG0 + (1 - corr)
        if (CI_TgG > 1) or (CI_TgG == np.inf):
            CI_TgG = 1
        Tg, Tgwall = flux.surface_temperature(f['TgK'], f['Tstart'], TgK_wall, Tstart_wall, f['TmaxLST'],
                                              TmaxLST_wall, altmax, dectime, SNUP, CI_TgG, landcover,
                                              shape=prep.shape, **opts)
        gr = groundview.gvf(prep.gvf, wallsun, shadow, Tg, Tgwall, Ta, f['emis_grid'], ewall, f['alb_grid'], SBC,
                            albedo_b, Twater=Twater, lc_grid=f['lc_grid'], landcover=landcover, **opts)
        (gvfLup, gvfalb, gvfalbnosh, gvfLupE, gvfalbE, gvfalbnoshE, gvfLupS, gvfalbS, gvfalbnoshS, gvfLupW,
         gvfalbW, gvfalbnoshW, gvfLupN, gvfalbN, gvfalbnoshN, _, _) = gr.outputs
        Tg = gr.surface_temperature
        Lup, _, Tgmap1 = flux.TsWaveDelay_2015a(gvfLup, firstdaytime, timeadd, timestepdec, Tgmap1, **opts)
        LupE, _, Tgmap1E = flux.TsWaveDelay_2015a(gvfLupE, firstdaytime, timeadd, timestepdec, Tgmap1E, **opts)
        LupS, _, Tgmap1S = flux.TsWaveDelay_2015a(gvfLupS, firstdaytime, timeadd, timestepdec, Tgmap1S, **opts)
        LupW, _, Tgmap1W = flux.TsWaveDelay_2015a(gvfLupW, firstdaytime, timeadd, timestepdec, Tgmap1W, **opts)
        LupN, _, Tgmap1N = flux.TsWaveDelay_2015a(gvfLupN, firstdaytime, timeadd, timestepdec, Tgmap1N, **opts)
        TgTemp = _terms(2, (Tg, shadow), (Ta,), prep)
        TgOut, timeadd, TgOut1 = flux.TsWaveDelay_2015a(TgTemp, firstdaytime, timeadd, timestepdec, TgOut1, **opts)
        F_sh = flux.cylindric_wedge(zen, f['svfalfa'], rows, cols, **opts)
        F_sh = _terms(4, (F_sh,), (0.5,), prep)
        Kdown = _terms(1, (shadow, dRad, f['svfbuveg'], F_sh),
                       (radI, np.sin(altitude * (np.pi / 180)), albedo_b, radG, radD), prep)
        Kup, KupE, KupS, KupW, KupN = flux.Kup_veg_2015a(radI, radD, radG, altitude, f['svfbuveg'], albedo_b, F_sh,
                                                         gvfalb, gvfalbE, gvfalbS, gvfalbW, gvfalbN, gvfalbnosh,
                                                         gvfalbnoshE, gvfalbnoshS, gvfalbnoshW, gvfalbnoshN, **opts)
        Keast, Ksouth, Kwest, Knorth, KsideI, KsideD, Kside = anisotropic.Kside_veg_v2022a(
            radI, radD, radG, shadow, f['svfS'], f['svfW'], f['svfN'], f['svfE'], f['svfEveg'], f['svfSveg'],
            f['svfWveg'], f['svfNveg'], azimuth, altitude, psi, t, albedo_b, F_sh, KupE, KupS, KupW, KupN, cyl, lv,
            1, None, rows, cols, f['asvf'], None, None, None, prepared=prep.shortwave, **opts)
        firstdaytime = 0
    else:
        Tgwall = 0.
        zero = _field(0.0, prep.shape, D)
        Kdown = Kwest = Kup = Keast = Ksouth = Knorth = KsideI = KsideD = F_sh = Tg = shadow = zero
        CI_Tg = CI_TgG = CI
        dRad = Kside = zero
        water = SBC * 0.98 * (float(Twater) + 273.15) ** 4 if landcover else 0.0
        Lup = _terms(3, (f['emis_grid'], f['lc_grid']), (Ta, SBC, int(landcover), water), prep, fp64_fields=(1,))
        LupE = LupS = LupW = LupN = Lup
        TgOut = _field(float(D.type(Ta)), prep.shape, D)
        I0 = 0
        timeadd = 0
        firstdaytime = 1
        lv = None
    L_patches = lv.copy() if altitude > 0 else prep.sky.copy()
    if CI < 0.95:
        esky = CI * esky + (1 - CI) * 1.
    Ldown, Lside, Least_, Lwest_, Lnorth_, Lsouth_ = anisotropic.longwave(
        prep.longwave if altitude > 0 else prep.night_longwave, esky, Ta, Tgwall, ewall, Lup, altitude, azimuth,
        f['asvf'], **opts)
    Least, Lsouth, Lwest, Lnorth = flux.Lside_veg_v2022a(
        f['svfS'], f['svfW'], f['svfN'], f['svfE'], f['svfEveg'], f['svfSveg'], f['svfWveg'], f['svfNveg'],
        f['svfEaveg'], f['svfSaveg'], f['svfWaveg'], f['svfNaveg'], azimuth, altitude, Ta, Tgwall, SBC, ewall, Ldown,
        esky, t, F_sh, CI, LupE, LupS, LupW, LupN, 1, **opts)
    if cyl == 0:
        Least, Lsouth, Lwest, Lnorth = (_add(a, b, prep) for a, b in
                                        ((Least, Least_), (Lsouth, Lsouth_), (Lwest, Lwest_), (Lnorth, Lnorth_)))
    Tmrt, _ = flux.assemble_tmrt(Kdown, Kup, Knorth, Keast, Ksouth, Kwest, Ldown, Lup, Lnorth, Least, Lsouth, Lwest,
                                 KsideI, Kside, Lside, absK, absL, Fside, Fup, Fcyl, cyl, 1, **opts)
    if cyl == 1:
        Least, Lsouth, Lwest, Lnorth = (_add(a, b, prep) for a, b in
                                        ((Least, Least_), (Lsouth, Lsouth_), (Lwest, Lwest_), (Lnorth, Lnorth_)))
    vals = (Tmrt, Kdown, Kup, Ldown, Lup, Tg, ea, esky, I0, CI, shadow, firstdaytime, timestepdec, timeadd, Tgmap1,
            Tgmap1E, Tgmap1S, Tgmap1W, Tgmap1N, Keast, Ksouth, Kwest, Knorth, Least, Lsouth, Lwest, Lnorth, KsideI,
            TgOut1, TgOut, radI, radD, Lside, L_patches, CI_Tg, CI_TgG, KsideD, dRad, Kside)
    return dict(zip(model.OUTPUT_NAMES, vals))


def as_float32(plane, prep):
    """A plane of the step rounded to float32 on the device (an FP32 plane is returned as it is)."""
    if plane.dtype == np.float32:
        return plane
    return _terms(6, (plane,), (), prep, fp64_fields=(0,), out_dtype=np.float32)


def _add(a, b, prep):
    """a + b in the working precision (the cardinal longwave sums of model.py:244-251)."""
    return _terms(5, (a, b), (), prep)
