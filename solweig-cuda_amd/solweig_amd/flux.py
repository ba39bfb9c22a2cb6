"""Per-step sub-models of the reference's model.py as standalone device calls (model.py:736-958).

cylindric_wedge, TsWaveDelay_2015a, Kup_veg_2015a, Lside_veg_v2022a, surface_temperature and
assemble_tmrt keep the reference's names, argument order, ``dtype`` keyword and return tuples;
they run as elementwise kernels of libsolweig_hip (sw_cylindric_wedge ... sw_assemble_tmrt) in
the requested precision. Inputs may be device arrays (runtime.DeviceArray / DeviceView), host
arrays or scalars; a single value is broadcast over the raster. Outputs are device views of one
[k, rows, cols] allocation. Perez_v3 is host FP64 code in the reference too (sky.perez).
The fused hour pass does not call these: they are the building blocks for callers of the
per-step API.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import runtime as rt
from . import sky

_SIGS = {}


def _entry(name):
    L = rt.lib()
    f = getattr(L, name)
    if name not in _SIGS:
        vp = ctypes.c_void_p
        f.argtypes = [vp, vp, ctypes.c_int, ctypes.c_longlong, vp, ctypes.c_int, ctypes.c_int, vp, vp]
        f.restype = ctypes.c_int
        _SIGS[name] = True
    return f


def _shape_of(values, shape):
    if shape is not None:
        return tuple(int(x) for x in shape)
    for v in values:
        s = getattr(v, 'shape', None)
        if s is not None and len(s) == 2 and int(np.prod(s)) > 1:
            return tuple(int(x) for x in s)
    return (1, 1)


def working_dtype(dtype):
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError('dtype must be float32 or float64')
    return dtype


def field_table(values, n, dtype, fp64_fields=()):
    """Device pointers and value counts of the fields of a sub-model call: device arrays of the
    expected dtype are passed as they are, anything else is uploaded (the uploads are returned
    and must stay alive until the call has finished)."""
    ptrs, counts, keep = [], [], []
    for k, v in enumerate(values):
        want = np.dtype(np.float64) if k in fp64_fields else dtype
        if isinstance(v, (rt.DeviceArray, rt.DeviceView)):
            if v.dtype != want:
                raise ValueError(f'field {k}: device array of {v.dtype}, expected {want}')
            c = int(np.prod(v.shape))
            ptrs.append(v.ptr)
        else:
            a = np.ascontiguousarray(np.asarray(v.get() if hasattr(v, 'get') else v, dtype=want)).reshape(-1)
            c = a.size
            d = rt.DeviceArray.from_host(a)
            keep.append(d)
            ptrs.append(d.ptr)
        if c not in (1, n):
            raise ValueError(f'field {k} holds {c} values; expected 1 or {n}')
        counts.append(c)
    return (ctypes.c_void_p * len(ptrs))(*ptrs), (ctypes.c_longlong * len(counts))(*counts), keep


def _run(name, values, params, n_out, dtype, shape=None, fp64_fields=()):
    dtype = working_dtype(dtype)
    shape = _shape_of(values, shape)
    n = int(np.prod(shape))
    P, C, keep = field_table(values, n, dtype, fp64_fields)
    out = rt.DeviceArray((n_out, *shape), dtype)
    Q = (ctypes.c_double * len(params))(*[float(x) for x in params])
    rt.check(_entry(name)(P, C, len(values), n, Q, len(params), dtype.itemsize, out.ptr, None))
    rt.check(rt.lib().sw_stream_sync(None))
    del keep
    return tuple(out.plane(k, shape) for k in range(n_out))


def cylindric_wedge(zen, svfalfa, rows, cols, *, dtype=np.float64):
    """F_sh, the sunlit fraction of walls seen by a cylinder (model.py:736-739)."""
    return _run('sw_cylindric_wedge', (svfalfa,), (float(zen),), 1, dtype, (int(rows), int(cols)),
                fp64_fields=(0,))[0]


def TsWaveDelay_2015a(gvfLup, firstdaytime, timeadd, timestepdec, Tgmap1, *, dtype=np.float64):
    """Surface temperature lag (model.py:742-756): (delayed, next timeadd, next state)."""
    timeadd, timestepdec = float(timeadd), float(timestepdec)
    update = timeadd >= 59 / 1440
    if update:
        weight = math.exp(-33.27 * timeadd)
        nxt = timestepdec if timestepdec > 59 / 1440 else 0.0
    else:
        nxt = timeadd + timestepdec
        weight = math.exp(-33.27 * nxt)
    d, s = _run('sw_wave_delay', (gvfLup, Tgmap1), (int(firstdaytime), int(update), weight), 2, dtype)
    return d, nxt, s


def Kup_veg_2015a(radI, radD, radG, altitude, svfbuveg, albedo_b, F_sh, gvfalb, gvfalbE, gvfalbS, gvfalbW,
                  gvfalbN, gvfalbnosh, gvfalbnoshE, gvfalbnoshS, gvfalbnoshW, gvfalbnoshN, *, dtype=np.float64):
    """Reflected shortwave from the ground (model.py:759-768): (Kup, KupE, KupS, KupW, KupN)."""
    return _run('sw_kup', (svfbuveg, F_sh, gvfalb, gvfalbE, gvfalbS, gvfalbW, gvfalbN, gvfalbnosh, gvfalbnoshE,
                           gvfalbnoshS, gvfalbnoshW, gvfalbnoshN),
                (radI, radD, radG, altitude, albedo_b), 5, dtype)


def Lside_veg_v2022a(svfS, svfW, svfN, svfE, svfEveg, svfSveg, svfWveg, svfNveg, svfEaveg, svfSaveg, svfWaveg,
                     svfNaveg, azimuth, altitude, Ta, Tw, SBC, ewall, Ldown, esky, t, F_sh, CI, LupE, LupS, LupW,
                     LupN, anisotropic_longwave, *, dtype=np.float64):
    """Longwave on the four cardinal faces (model.py:771-781): (Least, Lsouth, Lwest, Lnorth)."""
    return _run('sw_lside', (svfE, svfS, svfW, svfN, svfEveg, svfSveg, svfWveg, svfNveg, svfEaveg, svfSaveg,
                             svfWaveg, svfNaveg, Tw, Ldown, F_sh, LupE, LupS, LupW, LupN),
                (azimuth, altitude, Ta, SBC, ewall, esky, t, CI, int(anisotropic_longwave)), 4, dtype)


def surface_temperature(TgK, Tstart, TgK_wall, Tstart_wall, TmaxLST, TmaxLST_wall, altmax, dectime, SNUP,
                        CI_TgG, landcover=0, *, shape=None, dtype=np.float64):
    """Ground and wall temperature of the hour (model.py:784-793): (Tg, Tgwall)."""
    dectime = float(dectime)
    return _run('sw_surface_temperature', (TgK, Tstart, TgK_wall, Tstart_wall, TmaxLST, TmaxLST_wall),
                (altmax, dectime - math.floor(dectime), SNUP, CI_TgG, int(landcover)), 2, dtype, shape)


def assemble_tmrt(Kdown, Kup, Knorth, Keast, Ksouth, Kwest, Ldown, Lup, Lnorth, Least, Lsouth, Lwest, KsideI,
                  Kside, Lside, absK, absL, Fside, Fup, Fcyl, cyl, anisotropic_sky, *, SBC=5.67051e-8,
                  dtype=np.float64):
    """Mean radiant temperature (model.py:796-805): (Tmrt, Sstr)."""
    return _run('sw_assemble_tmrt', (Kdown, Kup, Knorth, Keast, Ksouth, Kwest, Ldown, Lup, Lnorth, Least, Lsouth,
                                     Lwest, KsideI, Kside, Lside),
                (absK, absL, Fside, Fup, Fcyl, int(cyl), int(anisotropic_sky), SBC), 2, dtype)


def Perez_v3(zen, azimuth, radD, radI, jday, patchchoice, patch_option):
    """Perez sky luminance (model.py:808-958, with its cosine clip; ``zen`` in degrees as the
    step passes it): (lv, clearness, brightness); lv is [153, 3] (altitude, azimuth, luminance)
    for patchchoice 1, else the luminance vector."""
    if int(patch_option) != 2:
        raise ValueError('only patch_option 2 (153 patches) is supported')
    alt, azi, lv, clear, bright = sky.perez(float(zen), float(azimuth), float(radD), float(radI), float(jday),
                                            clip=True, full=True)
    if int(patchchoice) == 1:
        lv = np.column_stack((alt, azi, lv))
    return lv, clear, bright
