"""Anisotropic-sky sub-models of one step as standalone device calls (model.py:259-506).

``prepare_geometry``, ``longwave``, ``prepare_shortwave`` and ``Kside_veg_v2022a`` keep the
reference's names, arguments and results. prepare_geometry packs the caller's three
[rows, cols, K] visibility masks into four flag bit planes on the device (sw_pack_visibility);
longwave and Kside_veg_v2022a build the per-patch FP64 coefficients of the call on the host and
run one pixel kernel each (sw_longwave, sw_kside) in the requested precision. Any patch set is
accepted (the sky vault of the night step, Perez luminances by day). The fused hour pass
computes the same sums from the domain's own visibility bits and does not call these.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import runtime as rt
from . import sky
from .flux import field_table, working_dtype

DEG = np.pi / 180
SBC = 5.67051e-8
COLS = 20
LW = dict(LDN=0, LSD=1, VEG_C=2, VEG_S=3, SH_C=4, SH_S=5, STER=6, COS=7, SIN=8, DE=9, DW=10, DN=11, DS=12, YI=13,
          THR_A=14, THR_B=15, ELIG=16)
KS = dict(LUMCHI=0, ANGC=1, STER=2, ANG=3, SUN_C=7, SH_C=8, DIRC=9, YI=13, THR_A=14, THR_B=15, FLAGS=16)
_SIGS = {}


def _entry(name, argtypes):
    f = getattr(rt.lib(), name)
    if name not in _SIGS:
        f.argtypes = argtypes
        f.restype = ctypes.c_int
        _SIGS[name] = True
    return f


def _host(v):
    if isinstance(v, (rt.DeviceArray, rt.DeviceView)):
        return v.to_host()
    return v.get() if hasattr(v, 'get') else v


@dataclass(frozen=True)
class PreparedGeometry:
    """A patch set and the packed visibility flags of a raster (model.py:259-306)."""
    sky_patches: np.ndarray
    steradian: np.ndarray
    directional_cosines: np.ndarray
    visibility: object
    shape: tuple

    @property
    def patch_count(self):
        return len(self.sky_patches)


@dataclass(frozen=True)
class PreparedShortwave:
    """PreparedGeometry plus the diffuse transmissivities as [K, P] (model.py:396-411)."""
    geometry: PreparedGeometry
    diffuse_visibility: object


def directional_cosines(azi):
    """[K, 4] cosines of the patch azimuth on the faces E, W, N, S, 0 outside a face's sector
    (model.py:298-303)."""
    return np.column_stack([
        np.where((azi > 360) | (azi < 180), np.cos((90 - azi) * DEG), 0.0),
        np.where((azi > 180) & (azi < 360), np.cos((270 - azi) * DEG), 0.0),
        np.where((azi > 270) | (azi < 90), np.cos((0 - azi) * DEG), 0.0),
        np.where((azi > 90) & (azi < 270), np.cos((180 - azi) * DEG), 0.0)])


def _mask_source(m, shape):
    """A [rows, cols, K] mask as (device pointer, value bytes, keep-alive)."""
    if isinstance(m, (rt.DeviceArray, rt.DeviceView)):
        if tuple(m.shape) != tuple(shape) or m.dtype.itemsize not in (1, 4, 8):
            raise ValueError('prepare_geometry: device masks must be [rows, cols, K] of 1, 4 or 8 byte values')
        return m.ptr, m.dtype.itemsize, None
    a = np.asarray(m.get() if hasattr(m, 'get') else m)
    if a.shape != tuple(shape):
        raise ValueError(f'prepare_geometry: mask of shape {a.shape}, expected {tuple(shape)}')
    if a.dtype == np.bool_:
        a = a.view(np.uint8)
    elif a.dtype.itemsize not in (1, 4, 8) or a.dtype.kind not in 'fu' or (a.dtype.kind == 'u' and a.dtype.itemsize != 1):
        a = a.astype(np.float64)
    d = rt.DeviceArray.from_host(np.ascontiguousarray(a))
    return d.ptr, a.dtype.itemsize, d


def prepare_geometry(sky_patches, shmat, vegshmat, vbshvegshmat):
    """model.py:273-306: ``sky_patches`` [K, >=2] (altitude, azimuth in degrees), masks
    [rows, cols, K] (host or device, float32/float64/uint8/bool; the three may differ)."""
    sky_ = np.array(_host(sky_patches), dtype=np.float64, copy=True)
    K = sky_.shape[0]
    shape = tuple(int(s) for s in np.shape(shmat)[:2])
    full = (*shape, K)
    srcs = [_mask_source(m, full) for m in (shmat, vegshmat, vbshvegshmat)]
    sizes = {s[1] for s in srcs}
    if len(sizes) != 1:                        # one value type for the pack kernel
        srcs = [_mask_source(np.asarray(_host(m), dtype=np.float64), full)
                for m in (shmat, vegshmat, vbshvegshmat)]
    n = shape[0] * shape[1]
    words = -(-K // 32)
    flags = rt.DeviceArray((4, words, n), np.uint32)
    vp = ctypes.c_void_p
    f = _entry('sw_pack_visibility', [vp, vp, vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, vp, vp])
    rt.check(f(srcs[0][0], srcs[1][0], srcs[2][0], srcs[0][1], n, K, flags.ptr, None))
    rt.check(rt.lib().sw_stream_sync(None))
    del srcs
    alt, azi = sky_[:, 0], sky_[:, 1]
    ster = sky.steradians(alt)
    dirs = directional_cosines(azi)
    for a in (sky_, ster, dirs):
        a.setflags(write=False)
    return PreparedGeometry(sky_, ster, dirs, flags, shape)


def _classification(alt, azi, solar_altitude, solar_azimuth):
    """Per patch: yi of shaded_or_sunlit (umep_solweig.py:1362-1397) clipped at 0 and the exact
    thresholds of tan(asvf) + yi (sunlit below the first, shaded from the second)."""
    xi = np.cos(np.abs(solar_azimuth - azi) * (np.pi / 180))
    yi = 2 * xi * np.tan(solar_altitude * (np.pi / 180))
    yi_ = np.where(yi > 0, 0.0, yi)
    thr = np.array([sky.sunlit_thresholds(float(a)) for a in alt]).reshape(-1, 2)
    return yi_, thr


def longwave_table(geometry, esky, Ta, ewall, solar_altitude, solar_azimuth):
    """[K, 20] FP64 coefficients of sw_longwave (Lcyl_v2022a's patch terms, model2's
    Martin-Berdahl band emissivities; umep_solweig.py:509-541, 946-1029)."""
    alt, azi = geometry.sky_patches[:, 0], geometry.sky_patches[:, 1]
    ster = geometry.steradian
    levels = np.unique(alt)
    band = 1 - (1 - esky) * np.exp(0.308 * (1.7 - (1 / np.cos((90 - levels) * DEG))))
    t = np.zeros((len(alt), COLS))
    for lvl, e in zip(levels, band):
        sel = alt == lvl
        flux = (e * SBC * ((Ta + 273.15) ** 4)) / np.pi
        t[sel, LW['LDN']] = flux * ster[sel] * np.sin(lvl * DEG)
        t[sel, LW['LSD']] = flux * ster[sel] * np.cos(lvl * DEG)
    surface = ((ewall * SBC * ((Ta + 273.15) ** 4)) / np.pi)
    cos_, sin_ = np.cos(alt * DEG), np.sin(alt * DEG)
    t[:, LW['VEG_C']] = surface * ster * cos_
    t[:, LW['VEG_S']] = surface * ster * sin_
    t[:, LW['SH_C']] = surface * ster * cos_
    t[:, LW['SH_S']] = surface * ster * sin_
    t[:, LW['STER']], t[:, LW['COS']], t[:, LW['SIN']] = ster, cos_, sin_
    t[:, LW['DE']:LW['DS'] + 1] = geometry.directional_cosines
    diff = np.abs(solar_azimuth - azi)
    t[:, LW['ELIG']] = ((diff > 90) & (diff < 270) & (solar_altitude > 0)).astype(np.float64)
    yi_, thr = _classification(alt, azi, solar_altitude, solar_azimuth)
    t[:, LW['YI']], t[:, LW['THR_A']], t[:, LW['THR_B']] = yi_, thr[:, 0], thr[:, 1]
    return t


def _tan_asvf(asvf, shape):
    a = np.asarray(_host(asvf), dtype=np.float64)
    return np.tan(a if a.size == 1 else a.reshape(shape))


def longwave(geometry, esky, Ta, Tgwall, ewall, Lup, solar_altitude, solar_azimuth, asvf, *, dtype=np.float64):
    """Anisotropic longwave of one step (model.py:367-392) -> (Ldown, Lside, Least, Lwest,
    Lnorth, Lsouth) device planes."""
    g = geometry
    dtype = working_dtype(dtype)
    esky, Ta, ewall, salt, sazi = (float(v) for v in (esky, Ta, ewall, solar_altitude, solar_azimuth))
    n = g.shape[0] * g.shape[1]
    tab = rt.DeviceArray.from_host(longwave_table(g, esky, Ta, ewall, salt, sazi))
    P, C, keep = field_table((Lup, Tgwall, _tan_asvf(asvf, g.shape)), n, dtype, fp64_fields=(2,))
    out = rt.DeviceArray((6, *g.shape), dtype)
    Q = (ctypes.c_double * 3)(Ta, ewall, SBC)
    vp, ci = ctypes.c_void_p, ctypes.c_int
    f = _entry('sw_longwave', [vp, ci, ctypes.c_longlong, vp, vp, vp, ci, vp, ci, ci, vp, vp])
    rt.check(f(g.visibility.ptr, g.patch_count, n, tab.ptr, P, C, 3, Q, 3, dtype.itemsize, out.ptr, None))
    rt.check(rt.lib().sw_stream_sync(None))
    del keep, tab
    return tuple(out.plane(k, g.shape) for k in range(6))


def prepare_shortwave(geometry, diffsh, *, dtype=np.float64):
    """model.py:401-411: the [rows, cols, K] diffuse transmissivities as [K, P] device values."""
    dtype = working_dtype(dtype)
    d = np.asarray(_host(diffsh))
    n, K = geometry.shape[0] * geometry.shape[1], geometry.patch_count
    if d.shape != (*geometry.shape, K):
        raise ValueError(f'prepare_shortwave: diffsh of shape {d.shape}, expected {(*geometry.shape, K)}')
    t = np.ascontiguousarray(d.reshape(n, K).T, dtype=dtype)
    return PreparedShortwave(geometry, rt.DeviceArray.from_host(t))


def shortwave_table(geometry, lv, radD, radI, altitude, azimuth, t, albedo):
    """[K, 20] FP64 coefficients of sw_kside (Kside_veg_v2022a's anisotropic patch terms,
    umep_solweig.py:747-905)."""
    alt, azi, lum = lv[:, 0], lv[:, 1], lv[:, 2]
    ster = geometry.steradian
    rad_tot = np.zeros(1)
    for k in range(len(alt)):
        rad_tot += lum[k] * ster[k] * np.sin(alt[k] * DEG)
    tab = np.zeros((len(alt), COLS))
    tab[:, KS['LUMCHI']] = (lum * radD) / rad_tot
    cos_ = np.cos(alt * DEG)
    tab[:, KS['ANGC']] = cos_ * np.cos(0)
    tab[:, KS['STER']] = ster
    normals = (90, 180, 270, 0)                        # faces E, S, W, N
    diff_sec = ((azi > 360) | (azi <= 180), (azi > 90) & (azi <= 270), (azi > 180) & (azi <= 360),
                (azi > 270) | (azi <= 90))
    refl_sec = ((azi > 360) | (azi < 180), (azi > 90) & (azi < 270), (azi > 180) & (azi < 360),
                (azi > 270) | (azi < 90))
    flags = np.zeros(len(alt), dtype=np.int64)
    for f, nrm in enumerate(normals):
        c = np.cos((nrm - azi + t) * DEG)
        tab[:, KS['ANG'] + f] = cos_ * c
        tab[:, KS['DIRC'] + f] = c
        flags |= diff_sec[f].astype(np.int64) << f
        flags |= refl_sec[f].astype(np.int64) << (4 + f)
    d = np.abs(azimuth - azi)
    flags |= ((d > 90) & (d < 270)).astype(np.int64) << 8
    sunlit = ((albedo * (radI * np.cos(altitude * DEG)) + (radD * 0.5)) / np.pi)
    shaded = ((albedo * radD * 0.5) / np.pi)
    tab[:, KS['SUN_C']] = sunlit * ster * cos_
    tab[:, KS['SH_C']] = shaded * ster * cos_
    yi_, thr = _classification(alt, azi, altitude, azimuth)
    tab[:, KS['YI']], tab[:, KS['THR_A']], tab[:, KS['THR_B']] = yi_, thr[:, 0], thr[:, 1]
    tab[:, KS['FLAGS']] = flags
    return tab


def Kside_veg_v2022a(radI, radD, radG, shadow, svfS, svfW, svfN, svfE, svfEveg, svfSveg, svfWveg, svfNveg, azimuth,
                     altitude, psi, t, albedo, F_sh, KupE, KupS, KupW, KupN, cyl, lv, anisotropic_diffuse, diffsh, rows,
                     cols, asvf, shmat, vegshmat, vbshvegshmat, *, dtype=np.float64, prepared=None):
    """Shortwave on the body's sides (model.py:443-506) -> (Keast, Ksouth, Kwest, Knorth, KsideI,
    KsideD, Kside) device planes; every (cyl, anisotropic_diffuse) branch."""
    dtype = working_dtype(dtype)
    shape = (int(rows), int(cols))
    n = shape[0] * shape[1]
    radI, radD, radG, azimuth, altitude, psi, t, albedo = (
        float(v) for v in (radI, radD, radG, azimuth, altitude, psi, t, albedo))
    cyl, ani = int(float(cyl)), int(float(anisotropic_diffuse))
    keep = []
    flags = diff = tab = None
    patches = 0
    tan_asvf = 0.0
    if ani == 1:
        if prepared is None:
            geometry = prepare_geometry(lv, shmat, vegshmat, vbshvegshmat)
            prepared = prepare_shortwave(geometry, diffsh, dtype=dtype)
        g = prepared.geometry
        if prepared.diffuse_visibility.dtype != dtype:
            raise ValueError('Kside_veg_v2022a: prepared shortwave state of another dtype')
        lv_ = np.asarray(_host(lv), dtype=np.float64)
        tab = rt.DeviceArray.from_host(shortwave_table(g, lv_, radD, radI, altitude, azimuth, t, albedo))
        keep.append(tab)
        flags, diff, patches = g.visibility.ptr, prepared.diffuse_visibility.ptr, g.patch_count
        tan_asvf = _tan_asvf(asvf, shape)
    rad = np.pi / 180
    lit = (azimuth > 360 - t or azimuth <= 180 - t, 90 - t < azimuth <= 270 - t, 180 - t < azimuth <= 360 - t,
           azimuth <= 90 - t or azimuth > 270 - t)
    sines = [np.sin((azimuth - r + t) * rad) for r in (0, 90, 180, 270)]
    params = [radI, radD, radG, np.cos(altitude * rad), psi, albedo, *sines, *(float(x) for x in lit),
              float(cyl), float(ani)]
    values = (shadow, F_sh, KupE, KupS, KupW, KupN, svfE, svfS, svfW, svfN, svfEveg, svfSveg, svfWveg, svfNveg,
              tan_asvf)
    P, C, up = field_table(values, n, dtype, fp64_fields=(14,))
    out = rt.DeviceArray((7, *shape), dtype)
    Q = (ctypes.c_double * 16)(*params)
    vp, ci = ctypes.c_void_p, ctypes.c_int
    f = _entry('sw_kside', [vp, vp, ci, ctypes.c_longlong, vp, vp, vp, ci, vp, ci, ci, vp, vp])
    rt.check(f(flags, diff, patches, n, tab.ptr if tab is not None else None, P, C, 15, Q, 16, dtype.itemsize,
               out.ptr, None))
    rt.check(rt.lib().sw_stream_sync(None))
    del keep, up
    return tuple(out.plane(k, shape) for k in range(7))
