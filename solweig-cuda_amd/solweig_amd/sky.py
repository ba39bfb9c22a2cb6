"""Per-hour host tables of the radiation path (FP64 NumPy, host side of the plain and batched forms).

Everything here is a scalar or per-patch quantity; the per-pixel work runs on the device. The
scalar recursion over a day (clearness index carried into the night, wave-delay timing) follows
Solweig_2022a_calc (model.py:73-218; umep_solweig.py:116-291) and series._run_day
(series.py:396-449); the per-patch tables fold the anisotropic sky models of
Kside_veg_v2022a (umep_solweig.py:666-779), Lcyl_v2022a/model2 (509-541, 946-1029) and
define_patch_characteristics (1205-1359) into one row of numbers per patch.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from .runtime import (H, PD, PF, SW_DIFF_DIGITS, SW_DIFF_K, SW_HOUR_STRIDE, SW_PATCH_D, SW_PATCH_F, SW_PATCHES,
                      SW_VIS_WORDS, SWS_CSH, SWS_CSUN, SWS_ELIG, SWS_MASK, SWS_NIB, SW_SEARCH_WORDS, diff_rows)
from . import utci as _utci

SBC = 5.67051e-8
DEG = np.pi / 180
BAND_COUNTS = (31, 30, 28, 24, 19, 13, 7, 1)
BAND_ALTITUDES = (6, 18, 30, 42, 54, 66, 78, 90)
BAND_STARTS = (0, 4, 2, 5, 8, 0, 10, 0)
FIRSTDAYLEAF, LASTDAYLEAF = 97, 300
TG_TMAX = 15.0      # TmaxLST whose surface-temperature sine the hour record carries

# Standing person, wall and human constants of series.py:35-51.
CONSTANTS = dict(albedo_b=0.2, albedo_g=0.15, ewall=0.9, eground=0.95, absK=0.7, absL=0.95,
                 Fside=0.22, Fup=0.06, Fcyl=0.28, elvis=0.0, trans=0.03)


# ----------------------------------------------------------------------------- sky vault

@lru_cache(maxsize=1)
def _patch_centres():
    alt, azi = [], []
    for n, a, s in zip(BAND_COUNTS, BAND_ALTITUDES, BAND_STARTS):
        step = 360 / n
        for k in range(n):
            alt.append(float(a))
            azi.append(k * step + s)
    return np.array(alt), np.array(azi)


def patches():
    """153 patch centres (altitude, azimuth) in degrees (create_patches option 2)."""
    alt, azi = _patch_centres()
    return alt.copy(), azi.copy()


def steradians(alt):
    """Solid angle of each patch (umep_solweig.py:749-759)."""
    levels, counts = np.unique(alt, return_counts=True)
    out = np.zeros(alt.shape[0])
    for i in range(alt.shape[0]):
        c = counts[levels == alt[i]][0]
        if c > 1:
            out[i] = ((360 / c) * DEG) * (np.sin((alt[i] + alt[0]) * DEG) - np.sin((alt[i] - alt[0]) * DEG))
        else:
            out[i] = ((360 / c) * DEG) * (np.sin(alt[i] * DEG) - np.sin((alt[i - 1] + alt[0]) * DEG))
    return out


@lru_cache(maxsize=64)
def sunlit_thresholds(alt):
    """Exact split points of the wall-patch classification for one patch altitude.

    shaded_or_sunlit (umep_solweig.py:1362-1397) calls a wall patch sunlit when
    arctan(x) * 180/pi < alt and shaded when it is > alt, x = tan(asvf) + min(yi, 0). With this
    host's NumPy arctan, returns (a, b): sunlit iff x < a, shaded iff x >= b (a == b unless some
    x rounds to exactly alt). Found by bisection over the FP64 bit patterns, evaluating arctan on
    vectors as the array code does.
    """
    alt = float(alt)

    def deg(xs):
        v = np.arctan(np.asarray(xs, dtype=np.float64)) * (180 / np.pi)
        return v

    def at(bits):
        return float(deg(np.full(16, np.int64(bits)).view(np.float64))[0])

    zero_bits = 0                                           # x = +0.0
    inf_bits = int(np.float64(np.inf).view(np.int64))       # x = +inf

    def first_true(pred):
        # smallest non-negative x (by bit pattern, monotone for x >= 0) with pred(deg(x)) true
        a, b = zero_bits, inf_bits
        if pred(at(a)):
            return 0.0
        if not pred(at(b)):
            return float('inf')
        while b - a > 1:
            m = (a + b) // 2
            if pred(at(m)):
                b = m
            else:
                a = m
        return float(np.int64(b).view(np.float64))

    sun_below = first_true(lambda v: not (v < alt))
    shade_from = first_true(lambda v: v > alt)
    return sun_below, shade_from


# ----------------------------------------------------------------------------- host scalars

def daylen(doy, lat):
    """Sunrise hour SNUP (umep_solweig.py:430-448)."""
    dec = -23.45 * np.cos(2.0 * np.pi * (doy + 10.0) / 365.0)
    soc = np.tan(DEG * dec) * np.tan(DEG * lat)
    soc = min(max(soc, -1.0), 1.0)
    dayl = 12.0 + 24.0 * np.arcsin(soc) / np.pi
    return 12.0 - dayl / 2.0


_G = ((10., (3.37, 2.85, 2.80, 2.64)), (20., (2.99, 3.02, 2.70, 2.93)), (30., (3.60, 3.00, 2.98, 2.93)),
      (40., (3.04, 3.11, 2.92, 2.94)), (50., (2.70, 2.95, 2.77, 2.71)), (60., (2.52, 3.07, 2.67, 2.93)),
      (70., (1.76, 2.69, 2.61, 2.61)), (80., (1.60, 1.67, 2.24, 2.63)), (90., (1.11, 1.44, 1.94, 2.02)))


def clearness(zen, jday, Ta, RH, radG, lat, P):
    """(I0, CI, Kt) after Crawford & Duchon with the low-sun correction
    (clearnessindex_2013b, umep_solweig.py:1654-1734); RH as a fraction."""
    p = 1013. if P == -999.0 else P * 10.
    b = 2. * np.pi * jday / 365.
    D = np.sqrt((1.00011 + np.dot(0.034221, np.cos(b)) + np.dot(0.001280, np.sin(b))
                 + np.dot(0.000719, np.cos((2. * b))) + np.dot(0.000077, np.sin((2. * b)))))
    m = 35. * np.cos(zen) * ((1224. * (np.cos(zen) ** 2) + 1) ** (-1 / 2.))
    trpg = 1.021 - 0.084 * (m * (0.000949 * p + 0.051)) ** 0.5
    G = next(g for upper, g in _G if lat < upper)
    season = 0 if (jday > 335 or jday <= 60) else 1 if jday <= 152 else 2 if jday <= 244 else 3
    G = G[season]
    a2, b2 = 17.27, 237.7
    td = (b2 * (((a2 * Ta) / (b2 + Ta)) + np.log(RH))) / (a2 - (((a2 * Ta) / (b2 + Ta)) + np.log(RH)))
    td = (td * 1.8) + 32
    u = np.exp(0.1133 - np.log(G + 1) + 0.0393 * td)
    tw = 1 - 0.077 * ((u * m) ** 0.3)
    tar = 0.935 ** m
    I0 = 1370.0 * np.cos(zen) * trpg * tw * D * tar
    if abs(zen) > np.pi / 2:
        I0 = 0
    corr = 0.1473 * np.log(90 - (zen / np.pi * 180)) + 0.3454
    with np.errstate(divide='ignore', invalid='ignore'):
        CI = np.float64(radG) / I0 + (1 - corr)
    Kt = radG / (1370.0 * np.cos(zen) * D)
    if math.isnan(CI):
        CI = float('inf')
    return I0, CI, Kt


def _reindl(radG, altitude, Kt, Ta, RH):
    """Reindl et al. (1990) split before its limits: (radI, radD, clearness-index bin)."""
    s = np.sin(altitude * (np.pi / 180))
    k = 0 if Kt <= 0.3 else (1 if Kt < 0.78 else 2)
    if Ta <= -999.00 or RH <= -999.00 or np.isnan(Ta) or np.isnan(RH):
        radD = radG * ((1.020 - 0.248 * Kt) if k == 0 else (1.45 - 1.67 * Kt) if k == 1 else 0.147)
    else:
        rh = RH / 100
        if k == 0:
            radD = radG * (1 - 0.232 * Kt + 0.0239 * s - 0.000682 * Ta + 0.0195 * rh)
        elif k == 1:
            radD = radG * (1.329 - 1.716 * Kt + 0.267 * s - 0.00357 * Ta + 0.106 * rh)
        else:
            radD = radG * (0.426 * Kt - 0.256 * s + 0.00349 * Ta + 0.0734 * rh)
    return (radG - radD) / s, radD, k


def diffuse_fraction(radG, altitude, Kt, Ta, RH):
    """Reindl et al. (1990) split (diffusefraction, umep_solweig.py:1789-1832); RH in %."""
    radI, radD, _ = _reindl(radG, altitude, Kt, Ta, RH)
    if radI < 0:
        radI = 0
    if altitude < 1 and radI > radG:
        radI = radG
    if radD > radG:
        radD = radG
    return radI, radD


def hour_branches(alt, azi, zen, jday, Ta, RH, radG, P, latitude, CI=1.0):
    """The branches that one hour's host chain takes for its forcing: at night (0, CI < 0.95);
    by day (1, CI clamped, Reindl bin and its three limits, Perez clearness bin, CI_Tg and
    CI_TgG clamped, CI < 0.95). The chain is smooth in the forcing while these stay the same, so
    a central difference is valid only between points with equal branches."""
    if alt <= 0:
        return (0, CI < 0.95)
    I0, CI, Kt = clearness(zen, jday, Ta, RH / 100., radG, latitude, P)
    ci_clamp = bool(CI > 1 or CI == np.inf)
    CI = 1 if ci_clamp else CI
    rI, rD, kbin = _reindl(radG, alt, Kt, Ta, RH)
    limits = (rI < 0, alt < 1 and max(rI, 0) > radG, rD > radG)
    radI, radD = diffuse_fraction(radG, alt, Kt, Ta, RH)
    z = zen * (180 / np.pi) * DEG                     # perez's zenith, as it forms it
    clear = ((radD + radI) / (radD + 1.041 * np.power(z, 3))) / (1 + 1.041 * np.power(z, 3))
    pbin = int(np.searchsorted(_CLEAR_EDGES, clear))
    radI0, radD0 = diffuse_fraction(I0, alt, 1., Ta, RH)
    corr = 0.1473 * np.log(90 - (zen / np.pi * 180)) + 0.3454
    citg = (radG / radI0) + (1 - corr)
    citgg = (radG / (radI0 * np.sin(alt * DEG) + radD0)) + (1 - corr)
    return (1, ci_clamp, kbin, limits, pbin, bool(citg > 1 or citg == np.inf), bool(citgg > 1 or citgg == np.inf),
            CI < 0.95)


_PEREZ = np.array([  # rows a1..a4, b1..b4, c1..c4, d1..d4, e1..e4; columns = 8 clearness bins
    [1.3525, -1.2219, -1.1000, -0.5484, -0.6000, -1.0156, -1.0000, -1.0500],
    [-0.2576, -0.7730, -0.2515, -0.6654, -0.3566, -0.3670, 0.0211, 0.0289],
    [-0.2690, 1.4148, 0.8952, -0.2672, -2.5000, 1.0078, 0.5025, 0.4260],
    [-1.4366, 1.1016, 0.0156, 0.7117, 2.3250, 1.4051, -0.5119, 0.3590],
    [-0.7670, -0.2054, 0.2782, 0.7234, 0.2937, 0.2875, -0.3000, -0.3250],
    [0.0007, 0.0367, -0.1812, -0.6219, 0.0496, -0.5328, 0.1922, 0.1156],
    [1.2734, -3.9128, -4.5000, -5.6812, -5.6812, -3.8500, 0.7023, 0.7781],
    [-0.1233, 0.9156, 1.1766, 2.6297, 1.8415, 3.3750, -1.6317, 0.0025],
    [2.8000, 6.9750, 24.7219, 33.3389, 21.0000, 14.0000, 19.0000, 31.0625],
    [0.6004, 0.1774, -13.0812, -18.3000, -4.7656, -0.9999, -5.0000, -14.5000],
    [1.2375, 6.4477, -37.7000, -62.2500, -21.5906, -7.1406, 1.2438, -46.1148],
    [1.0000, -0.1239, 34.8438, 52.0781, 7.2492, 7.5469, -1.9094, 55.3750],
    [1.8734, -1.5798, -5.0000, -3.5000, -3.5000, -3.4000, -4.0000, -7.2312],
    [0.6297, -0.5081, 1.5218, 0.0016, -0.1554, -0.1078, 0.0250, 0.4050],
    [0.9738, -1.7812, 3.9229, 1.1477, 1.4062, -1.0750, 0.3844, 13.3500],
    [0.2809, 0.1080, -2.6204, 0.1062, 0.3988, 1.5702, 0.2656, 0.6234],
    [0.0356, 0.2624, -0.0156, 0.4659, 0.0032, -0.0672, 1.0468, 1.5000],
    [-0.1246, 0.0672, 0.1597, -0.3296, 0.0766, 0.4016, -0.3788, -0.6426],
    [-0.5718, -0.2190, 0.4199, -0.0876, -0.0656, 0.3017, -2.4517, 1.8564],
    [0.9938, -0.4285, -0.5562, -0.0329, -0.1294, -0.4844, 1.4656, 0.5636]])
_CLEAR_EDGES = (1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200)


def perez(zen_deg, azimuth, radD, radI, jday, clip=False, full=False):
    """Perez et al. (1993) relative luminance over the 153 patches (Perez_v3,
    umep_solweig.py:1836-2057): returns (alt, azi, lv) with the patch angles round-tripped
    deg -> rad -> deg as the original returns them; ``full`` also returns the sky clearness and
    brightness. ``clip`` limits the cosine of the patch-sun angle to [-1, 1] (model.py:933; the
    UMEP step does not clip)."""
    zen = zen_deg * DEG
    az = azimuth * DEG
    altitude = (90 - zen_deg) * DEG
    clear = ((radD + radI) / (radD + 1.041 * np.power(zen, 3))) / (1 + 1.041 * np.power(zen, 3))
    day_angle = jday * 2 * np.pi / 365
    I0 = 1367 * (1.00011 + 0.034221 * np.cos(day_angle) + 0.00128 * np.sin(day_angle)
                 + 0.000719 * np.cos(2 * day_angle) + 0.000077 * np.sin(2 * day_angle))
    if altitude >= 10 * DEG:
        airmass = 1 / np.sin(altitude)
    else:
        airmass = 1 / np.sin(altitude) + 0.50572 * np.power(180 * altitude / np.pi + 6.07995, -1.6364)
    bright = 0 if radD <= 10 else (airmass * radD) / I0
    k = None
    if clear < _CLEAR_EDGES[0]:
        k = 0
    for j in range(6):
        if _CLEAR_EDGES[j] < clear < _CLEAR_EDGES[j + 1]:
            k = j + 1
    if clear > _CLEAR_EDGES[6]:
        k = 7
    if k is None:
        raise ValueError(f'Perez clearness {clear} sits exactly on a bin edge')
    c = lambda row, col: _PEREZ[row * 4 + col, k]
    ma = c(0, 0) + c(0, 1) * zen + bright * (c(0, 2) + c(0, 3) * zen)
    mb = c(1, 0) + c(1, 1) * zen + bright * (c(1, 2) + c(1, 3) * zen)
    me = c(4, 0) + c(4, 1) * zen + bright * (c(4, 2) + c(4, 3) * zen)
    if k > 0:
        mc = c(2, 0) + c(2, 1) * zen + bright * (c(2, 2) + c(2, 3) * zen)
        md = c(3, 0) + c(3, 1) * zen + bright * (c(3, 2) + c(3, 3) * zen)
    else:
        mc = np.exp(np.power(bright * (c(2, 0) + c(2, 1) * zen), c(2, 2))) - 1
        md = -np.exp(bright * (c(3, 0) + c(3, 1) * zen)) + c(3, 2) + bright * c(3, 3) * bright
    alt, azi = patches()
    vzen, valt, vazi = (90 - alt) * DEG, alt * DEG, azi * DEG
    cosg = np.sin(valt) * np.sin(altitude) + np.cos(altitude) * np.cos(valt) * np.cos(np.abs(vazi - az))
    if clip:
        cosg = np.clip(cosg, -1.0, 1.0)
    lv = (1 + ma * np.exp(mb / np.cos(vzen))) * ((1 + mc * np.exp(md * np.arccos(cosg)) + me * cosg * cosg))
    lv = lv / np.sum(lv)
    if full:
        return valt * (180 / np.pi), vazi * (180 / np.pi), lv, clear, bright
    return valt * (180 / np.pi), vazi * (180 / np.pi), lv


def leaf_psi(doy):
    """1 - canopy shortwave transmissivity (series.py:416-422)."""
    return 3. / 100. if (doy > FIRSTDAYLEAF) and (doy < LASTDAYLEAF) else .5


def diff_veg(trans=0.03):
    """diffsh of a patch seen through canopy, FP32 (geometry.cu:161-172): 1 - (1 - 0)(1 - t)."""
    one = np.float32(1.0)
    return float(one - (one - np.float32(0.0)) * (one - np.float32(trans)))


# ----------------------------------------------------------------------------- ray tables

def shadow_ray(azimuth, altitude, scale, amaxvalue, shape, first=0):
    """Sun-ray offsets and heights of one hour (_shadow_ray, model.py:655-686, the loop of
    umep_solweig.py:2286-2358): offsets int32 [L, 2] (row, col), dz FP64 [L], facesh bounds."""
    azi = azimuth * (np.pi / 180.)
    alt = altitude * (np.pi / 180.)
    s, c, t = np.sin(azi), np.cos(azi), np.tan(azi)
    ymajor = np.pi / 4 <= azi < 3 * np.pi / 4 or 5 * np.pi / 4 <= azi < 7 * np.pi / 4
    ds = abs(1 / (s if ymajor else c))
    tanalt = np.tan(alt) / scale
    index = np.arange(first, max(shape) + 2, dtype=np.float64)
    if ymajor:
        dy = np.sign(s) * index
        dx = -np.sign(c) * np.abs(np.round(index / t))
    else:
        dy = np.sign(s) * np.abs(np.round(index * t))
        dx = -np.sign(c) * index
    dz = (ds * index) * tanalt
    bad = (~(amaxvalue >= dz)) | (np.abs(dx) >= shape[0]) | (np.abs(dy) >= shape[1])
    n = 0 if amaxvalue < 0 else int(np.flatnonzero(bad)[0]) + 1
    off = np.column_stack((dx[:n], dy[:n])).astype(np.int32)
    low, high = azi - np.pi / 2, azi + np.pi / 2
    middle = low >= 0 and high < 2 * np.pi
    if not middle:
        if low < 0:
            low += 2 * np.pi
        else:
            high -= 2 * np.pi
    return off, dz[:n].copy(), low, high, middle


STEP_BLOCK = 8   # the device reads ray and ground-view steps 8 at a time


def pad_steps(off, dz):
    """Ray offsets/heights padded to a multiple of STEP_BLOCK by repeating the last step (the
    hour record keeps the true length; padded steps are loaded but never evaluated)."""
    n = len(dz)
    m = -(-n // STEP_BLOCK) * STEP_BLOCK
    if m == n:
        return np.ascontiguousarray(off, np.int32), np.ascontiguousarray(dz, np.float64)
    o = np.empty((m, 2), np.int32)
    z = np.empty(m, np.float64)
    o[:n], z[:n] = off, dz
    o[n:], z[n:] = off[n - 1], dz[n - 1]
    return o, z


def gvf_geometry(scale=1.0, first=1.0, second=22.0):
    """18 search directions x second_steps offsets and facesh bounds (prepare_gvf_geometry,
    model.py:533-566; sunonsurface_2018a, umep_solweig.py:1466-1563).

    The step axis of ``offsets`` is padded to a multiple of 8 by repeating the last step (the
    device loads the steps of a direction 8 at a time); steps past ``second_steps`` are never
    summed."""
    first_steps = max(1, int(np.round(first * scale)))
    second_steps = int(np.round(second * scale))
    stride = -(-second_steps // 8) * 8
    offsets = np.empty((18, stride, 2), dtype=np.int32)
    bounds = np.empty((18, 3))
    for j, degrees in enumerate(np.arange(5, 359, 20)):
        az = degrees * (np.pi / 180)
        sa, ca, ta = np.sin(az), np.cos(az), np.tan(az)
        ydom = (np.pi / 4 <= az < 3 * np.pi / 4) or (5 * np.pi / 4 <= az < 7 * np.pi / 4)
        for n in range(second_steps):
            if ydom:
                dy, dx = np.sign(sa) * n, -np.sign(ca) * abs(np.round(n / ta))
            else:
                dy, dx = np.sign(sa) * abs(round(n * ta)), -np.sign(ca) * n
            offsets[j, n] = int(dx), int(dy)
        offsets[j, second_steps:] = offsets[j, second_steps - 1]      # see pad_steps
        low, high = az - np.pi / 2, az + np.pi / 2
        if low >= 0 and high < 2 * np.pi:
            bounds[j] = low, high, 0
        elif low < 0 and high <= 2 * np.pi:
            bounds[j] = low + 2 * np.pi, high, 1
        else:
            bounds[j] = low, high - 2 * np.pi, 1
    return offsets, bounds, first_steps, second_steps


# ----------------------------------------------------------------------------- patch rows

@lru_cache(maxsize=4)
def _geometry(round_trip):
    alt, azi = patches()
    if round_trip:
        alt, azi = alt * DEG * (180 / np.pi), azi * DEG * (180 / np.pi)
    ster = steradians(alt)
    wcos = ster * np.cos(alt * DEG)
    wsin = ster * np.sin(alt * DEG)
    dirs = np.column_stack([
        np.where((azi > 360) | (azi < 180), np.cos((90 - azi) * DEG), 0.0),
        np.where((azi > 180) & (azi < 360), np.cos((270 - azi) * DEG), 0.0),
        np.where((azi > 270) | (azi < 90), np.cos((0 - azi) * DEG), 0.0),
        np.where((azi > 90) & (azi < 270), np.cos((180 - azi) * DEG), 0.0)])
    thr = np.array([sunlit_thresholds(a) for a in alt])
    # per patch sin(altitude) evaluated one value at a time, as the luminance normalisation of
    # Kside_veg_v2022a forms it (umep_solweig.py:689-691), and the band of every patch
    sin_alt = [float(np.sin(a * DEG)) for a in alt]
    levels = np.unique(alt)
    band = np.searchsorted(levels, alt)
    return alt, azi, ster, wcos, wsin, dirs, thr, sin_alt, levels, band


def patch_rows(day, esky, Ta, altitude, azimuth, lv=None, radD=0.0, f64=False):
    """One hour's patch table: float [153, SW_PATCH_F] and double [153, SW_PATCH_D]
    (``f64``: the float table left in FP64, for derivatives)."""
    alt, azi, ster, wcos, wsin, dirs, thr, sin_alt, levels, band = _geometry(bool(day))
    f = np.zeros((SW_PATCHES, SW_PATCH_F))
    d = np.zeros((SW_PATCHES, SW_PATCH_D))
    band_emis = 1 - (1 - esky) * np.exp(0.308 * (1.7 - (1 / np.cos((90 - levels) * DEG))))
    e = band_emis[band]
    flux = (e * SBC * ((Ta + 273.15) ** 4)) / np.pi
    f[:, PF['LDN']] = flux * ster * np.sin(alt * DEG)
    f[:, PF['LSD']] = flux * ster * np.cos(alt * DEG)
    f[:, PF['WCOS']] = wcos
    f[:, PF['WSIN']] = wsin
    f[:, PF['DE']], f[:, PF['DW']], f[:, PF['DN']], f[:, PF['DS']] = dirs.T
    d[:, PD['SUN_BELOW']], d[:, PD['SHADE_FROM']] = thr[:, 0], thr[:, 1]
    if day:
        # sequential FP64 sum in patch order (Python floats round as NumPy float64 scalars)
        rad_tot = 0.0
        for lvk, sk, snk in zip(lv.tolist(), ster.tolist(), sin_alt):
            rad_tot += lvk * sk * snk
        lum_chi = (lv * radD) / np.array([rad_tot])
        f[:, PF['LV']] = lv
        f[:, PF['KD']] = lum_chi * (np.cos(alt * DEG) * np.cos(0)) * ster
        diff = np.abs(azimuth - azi)
        f[:, PF['ELIGIBLE']] = ((diff > 90) & (diff < 270) & (altitude > 0)).astype(np.float64)
        yi = 2 * np.cos(diff * DEG) * np.tan(altitude * DEG)
        d[:, PD['YI']] = np.where(yi > 0, 0.0, yi)
    return (f if f64 else f.astype(np.float32)), d


_NIB_BITS = ((np.arange(16)[:, None] >> np.arange(4)[None, :]) & 1).astype(np.float64)   # [16][4]
_INF_BITS = int(np.float64(np.inf).view(np.int64))


def breakpoints(y, t):
    """Elementwise the smallest double c >= 0 with c + y >= t in FP64 (+inf when no finite c
    reaches t): for every x >= 0, x + y < t exactly when x < c, since rounding is monotonic, so
    that the device's wall-patch test tan(asvf) + yi < threshold becomes tan(asvf) < c. Searched
    over the bit patterns of x >= 0 (which order as the values do), evaluating the sum as the device
    does: from the estimate t - y a few neighbouring patterns settle almost every element; the rest
    are bisected."""
    y, t = np.broadcast_arrays(np.asarray(y, dtype=np.float64), np.asarray(t, dtype=np.float64))
    shape = y.shape
    y, t = y.ravel(), t.ravel()

    def reached(bits, yy, tt):
        with np.errstate(invalid='ignore', over='ignore'):
            return (bits.view(np.float64) + yy) >= tt

    with np.errstate(invalid='ignore', over='ignore'):
        est = t - y
    est = np.clip(np.where(np.isnan(est), 0.0, est), 0.0, np.inf)
    bits = est.view(np.int64).copy()
    for _ in range(4):                                    # down while the pattern below reaches t
        below = np.maximum(bits - 1, 0)
        bits = np.where((bits > 0) & reached(below, y, t), below, bits)
    for _ in range(4):                                    # up while this pattern does not
        bits = np.where(~reached(bits, y, t) & (bits < _INF_BITS), bits + 1, bits)
    hit = reached(bits, y, t)
    exact = hit & ((bits == 0) | ~reached(np.maximum(bits - 1, 0), y, t))
    never = (bits == _INF_BITS) & ~hit                    # no finite x reaches t
    c = np.where(never, np.inf, bits.view(np.float64))
    rest = ~(exact | never)
    if rest.any():
        c[rest] = _breakpoints_bisect(y[rest], t[rest])
    return c.reshape(shape)


def _breakpoints_bisect(y, t):
    """breakpoints by bisection over all bit patterns of x >= 0 (64 steps)."""
    def reached(bits):
        with np.errstate(invalid='ignore', over='ignore'):
            return (bits.view(np.float64) + y) >= t

    lo = np.zeros(y.shape, np.int64)                      # +0.0: not reached (else c = 0)
    hi = np.full(y.shape, _INF_BITS, np.int64)            # +inf: reached (else c = +inf)
    at0, atinf = reached(lo), reached(hi)
    for _ in range(64):
        mid = lo + (hi - lo) // 2
        r = reached(mid)
        hi = np.where(r, mid, hi)
        lo = np.where(r, lo, mid)
    c = hi.view(np.float64)
    return np.where(at0, 0.0, np.where(atinf, c, np.inf))
_BAND_BITS = np.left_shift(np.uint32(1), np.arange(32, dtype=np.uint32))


def search_row(f, d, _defer=False):
    """The patch search tables of one hour (SWS_* words of sw_domain.ptab_s) from its float
    patch rows ``f`` [153][SW_PATCH_F] and double rows ``d`` [153][SW_PATCH_D]:

    * nibble sums: for the 4 patches 4i..4i+3 and every 4-bit subset, the FP64 sums (rounded
      once to FP32) of luminance and lumChi cos omega, so that a visibility word's diffuse sums
      are 8 table reads;
    * per altitude band, in the ascending yi order of its patches, the breakpoints of the sunlit
      and the shaded test (``breakpoints``: tan(asvf) + yi < threshold exactly when
      tan(asvf) < c) and the band-local bit masks of the j first patches of that order:
      tan(asvf) + yi is non-decreasing in yi (rounding is monotonic), so the patches of a band
      below each exact threshold are the first j of the order, j found by bisection over the
      non-increasing breakpoints;
    * per band the eligibility bits of the sunlit-wall test."""
    out = np.zeros(SW_SEARCH_WORDS, np.uint32)
    w = np.zeros((156, 2))
    w[:SW_PATCHES, 0] = f[:, PF['LV']]
    w[:SW_PATCHES, 1] = f[:, PF['KD']]
    nib = np.einsum('ej,ijc->iec', _NIB_BITS, w.reshape(39, 4, 2)).astype(np.float32)
    out[SWS_NIB:SWS_NIB + 39 * 16 * 2] = nib.reshape(-1).view(np.uint32)
    ys = np.full((8, 32), np.inf)
    t_sun = np.zeros((8, 1))
    t_sh = np.zeros((8, 1))
    masks = np.zeros((8, 33), np.uint32)
    elig = np.zeros(8, np.uint32)
    yi_all = np.asarray(d[:, PD['YI']], dtype=np.float64)
    el_all = np.asarray(f[:, PF['ELIGIBLE']]) != 0
    lo = 0
    for b, n in enumerate(BAND_COUNTS):
        yi = yi_all[lo:lo + n]
        order = np.argsort(yi, kind='stable')
        ys[b, :n] = yi[order]
        t_sun[b] = d[lo, PD['SUN_BELOW']]
        t_sh[b] = d[lo, PD['SHADE_FROM']]
        bits = np.left_shift(np.uint32(1), order.astype(np.uint32))
        masks[b, 1:n + 1] = np.bitwise_or.accumulate(bits)
        masks[b, n + 1:] = masks[b, n]
        elig[b] = np.bitwise_or.reduce(np.where(el_all[lo:lo + n], _BAND_BITS[:n], np.uint32(0)))
        lo += n
    out[SWS_MASK:SWS_MASK + 264] = masks.reshape(-1)
    out[SWS_ELIG:SWS_ELIG + 8] = elig
    if _defer:                       # build_day fills the breakpoints of all hours at once
        return out, (ys, t_sun, t_sh)
    return _with_breakpoints(out[None], [(ys, t_sun, t_sh)])[0]


def _with_breakpoints(rows, parts):
    """Fill the breakpoint tables of search rows [k][SW_SEARCH_WORDS] from their (sorted yi,
    sunlit and shaded thresholds), in one vectorised bisection for all rows."""
    if not len(parts):
        return rows
    ys = np.stack([p[0] for p in parts])
    cs = breakpoints(np.stack([ys, ys]), np.stack([np.stack([p[1] for p in parts]),
                                                   np.stack([p[2] for p in parts])]))
    rows[:, SWS_CSUN:SWS_CSUN + 512] = cs[0].reshape(len(parts), -1).view(np.uint32)
    rows[:, SWS_CSH:SWS_CSH + 512] = cs[1].reshape(len(parts), -1).view(np.uint32)
    return rows


# ----------------------------------------------------------------------------- a day

def diffuse_table(hours, ptab_f):
    """The day's diffuse table (sw_domain.diffuse_tab, the input of sw_day_diffuse) from its hour
    records [n][SW_HOUR_STRIDE] and float patch rows [rows][153][SW_PATCH_F], as bytes.

    Per daytime hour h, row 2h holds the patch luminance and row 2h + 1 lumChi cos omega (the
    per-patch terms that umep_solweig.py:148-154 and shortwave.cu:155-157 weight by diffsh): at
    positions k < 153 for the sky mask, at 160 + k times the hour's vegetation diffsh (as FP32,
    the product in FP64, exact) for the vegetation-only mask. Each row is rounded to integers q
    on a power-of-two scale with |q| <= 2^27 (27 bits of the row's largest magnitude) and stored
    as four int8 digits q = d3 2^21 + d2 2^14 + d1 2^7 + d0 (d0..d2 in [0, 127]); a visibility
    mask times a row is then an exact integer sum on the device. Rows of night hours and rows
    past the day carry the scale -1 (not written by the device); a row holding a non-finite value
    carries the scale NaN."""
    n = hours.shape[0]
    rows = diff_rows(n)
    vals = np.zeros((rows, SW_DIFF_K))
    day = np.zeros(rows, bool)
    for h in range(n):
        if not hours[h, H['DAY']]:
            continue
        f = ptab_f[int(hours[h, H['PATCH_TABLE']])]
        dveg = float(np.float32(hours[h, H['DIFF_VEG']]))
        for v, col in ((0, PF['LV']), (1, PF['KD'])):
            x = np.asarray(f[:, col], dtype=np.float32).astype(np.float64)
            vals[2 * h + v, :SW_PATCHES] = x
            vals[2 * h + v, 160:160 + SW_PATCHES] = dveg * x
        day[2 * h:2 * h + 2] = True
    # a non-finite value makes every pixel's sum NaN, as 0 x NaN does in the reference's
    # diffsh-weighted sum: the row gets q = 0 and the scale NaN
    finite = np.isfinite(vals).all(axis=1)
    vals[~finite] = 0.0
    top = np.abs(vals).max(axis=1)
    e = np.frexp(top)[1]                                  # top < 2^e (0 for an all-zero row)
    scale = np.where(finite, np.ldexp(1.0, e - 27), np.nan)
    q = np.rint(vals / np.where(finite, scale, 1.0)[:, None]).astype(np.int64)
    digits = np.empty((SW_DIFF_DIGITS, rows, SW_DIFF_K), np.int8)
    for d in range(SW_DIFF_DIGITS - 1):
        digits[d] = q & 127
        q >>= 7
    digits[SW_DIFF_DIGITS - 1] = q
    scale = np.where(day, scale, -1.0)
    # the device's operand order: per 8-hour block, step of 64 positions and digit, the 64 lanes'
    # 16-byte fragments (lane c + 16 g: row 16 block + c, positions 64 step + 16 g .. + 15)
    frag = digits.reshape(SW_DIFF_DIGITS, rows // 16, 16, SW_DIFF_K // 64, 4, 16).transpose(1, 3, 0, 4, 2, 5)
    return np.concatenate([np.ascontiguousarray(frag).reshape(-1).view(np.uint8), scale.view(np.uint8)])


def diffuse_digits(table, n_hours):
    """The int8 digits [SW_DIFF_DIGITS][rows][SW_DIFF_K] and the row scales of a diffuse table."""
    rows = diff_rows(n_hours)
    nd = SW_DIFF_DIGITS * rows * SW_DIFF_K
    frag = np.asarray(table[:nd]).view(np.int8).reshape(rows // 16, SW_DIFF_K // 64, SW_DIFF_DIGITS, 4, 16, 16)
    digits = frag.transpose(2, 0, 4, 1, 3, 5).reshape(SW_DIFF_DIGITS, rows, SW_DIFF_K)
    return digits, np.asarray(table[nd:nd + 8 * rows]).view(np.float64)


def diffuse_values(table, n_hours):
    """The rows of a diffuse table as the integers they hold, [rows][SW_DIFF_K] int64, and the
    row scales (value = q x |scale|)."""
    digits, scale = diffuse_digits(table, n_hours)
    q = sum(digits[d].astype(np.int64) << (7 * d) for d in range(SW_DIFF_DIGITS))
    return q, scale


class DayTables:
    """Host tables of every hour of one day for one domain (the inputs of sw_hours_forward)."""

    def __init__(self, hours, ptab_f, ptab_d, ray_off, ray_dz, meta, ptab_s=None):
        self.hours, self.ptab_f, self.ptab_d = hours, ptab_f, ptab_d
        self.ray_off, self.ray_dz, self.meta = ray_off, ray_dz, meta
        self.ptab_s = ptab_s
        self._diffuse = None

    @property
    def n_hours(self):
        return self.hours.shape[0]

    @property
    def diffuse(self):
        """The day's diffuse table (diffuse_table), built on first use."""
        if getattr(self, '_diffuse', None) is None:
            self._diffuse = diffuse_table(self.hours, self.ptab_f)
        return self._diffuse


def hour_tables(alt, azi, zen, jday, dectime, altmax, Ta, RH, radG, P, Ws, doy, latitude, shape, scale,
                amaxvalue, state, *, wall_params=(0.37, -3.41, 15.0), constants=CONSTANTS, psi=None,
                f64=False, rows=True, ray=True):
    """One step's hour record, patch rows and sun ray (Solweig_2022a_calc's host part,
    model.py:73-218). ``state`` = (CI, firstdaytime, timeadd, timestepdec) before the step; the
    returned state is the one after it (the CI reset at midnight rows is the caller's).
    ``f64``: every value as a smooth FP64 function of the forcing (no float32 rounding of the
    UTCI inputs or the patch table), for the derivatives of the adjoint; ``rows=False`` skips the
    Perez luminance and the patch rows (returned as None), ``ray=False`` the sun ray (its record
    entries stay 0)."""
    CI, firstdaytime, timeadd, timestepdec = state
    c = constants
    tgk_w, tstart_w, tmax_w = wall_params
    h = np.zeros(SW_HOUR_STRIDE)
    psi = leaf_psi(doy) if psi is None else psi
    snup = daylen(jday, latitude)
    ea = 6.107 * 10 ** ((7.5 * Ta) / (237.3 + Ta)) * (RH / 100.)
    msteg = 46.5 * (ea / (Ta + 273.15))
    esky = (1 - (1 + msteg) * np.exp(-((1.2 + 3.0 * msteg) ** 0.5))) + c['elvis']
    radI = radD = 0.0
    tgwall = 0.0
    lv = None
    I0, citg, citgg = 0.0, CI, CI
    off = np.zeros((0, 2), np.int32)
    dz = np.zeros(0)
    if alt > 0:
        I0, CI, Kt = clearness(zen, jday, Ta, RH / 100., radG, latitude, P)
        if (CI > 1) or (CI == np.inf):
            CI = 1
        radI, radD = diffuse_fraction(radG, alt, Kt, Ta, RH)
        if rows:
            _, _, lv = perez(zen * (180 / np.pi), azi, radD, radI, jday)
        radI0, radD0 = diffuse_fraction(I0, alt, 1., Ta, RH)
        corr = 0.1473 * np.log(90 - (zen / np.pi * 180)) + 0.3454
        citg = (radG / radI0) + (1 - corr)
        if (citg > 1) or (citg == np.inf):
            citg = 1
        radG0 = radI0 * np.sin(alt * DEG) + radD0
        citgg = (radG / radG0) + (1 - corr)
        if (citgg > 1) or (citgg == np.inf):
            citgg = 1
        frac = dectime - np.floor(dectime)
        # the ground-temperature wave's sine for TmaxLST = 15 (the default maps and most land-cover
        # classes); the device evaluates other TmaxLST values itself
        h[H['TG_TMAX']] = TG_TMAX
        h[H['TG_SINE']] = np.sin(((frac - snup / 24) / (TG_TMAX / 24 - snup / 24)) * np.pi / 2)
        tgwall = (tgk_w * altmax + tstart_w) * np.sin(((frac - snup / 24) / (tmax_w / 24 - snup / 24)) * np.pi / 2)
        if tgwall < 0:
            tgwall = 0
        tgwall = tgwall * citgg
        # the five Lup delays use the incoming timing; TgOut's call advances it (model.py:154-163)
        first = 1 if firstdaytime == 1 else 0
        if timeadd >= 59 / 1440:
            update, weight = 1, math.exp(-33.27 * timeadd)
            timeadd = timestepdec if timestepdec > 59 / 1440 else 0.0
        else:
            update = 0
            timeadd = timeadd + timestepdec
            weight = math.exp(-33.27 * timeadd)
        firstdaytime = 0
        h[H['DAY']] = 1
        h[H['CITGG']], h[H['FRAC']] = citgg, frac
        h[H['LAG_WEIGHT']], h[H['LAG_FIRST']], h[H['LAG_UPDATE']] = weight, first, update
        if ray:
            off, dz, low, high, middle = shadow_ray(azi, alt, scale, amaxvalue, shape)
            h[H['FACE_LOW']], h[H['FACE_HIGH']], h[H['FACE_MIDDLE']] = low, high, float(middle)
            h[H['RAY_LEN']] = len(dz)
    else:
        timeadd = 0.
        firstdaytime = 1
    if CI < 0.95:
        esky = CI * esky + (1 - CI) * 1.
    h[H['ALTITUDE']], h[H['AZIMUTH']], h[H['TANZEN']] = alt, azi, np.tan(zen)
    h[H['TA']], h[H['RADG']], h[H['RADD']], h[H['RADI']], h[H['PSI']] = Ta, radG, radD, radI, psi
    h[H['SINALT']], h[H['COSALT']] = np.sin(alt * (np.pi / 180)), np.cos(alt * DEG)
    h[H['ALTMAX']], h[H['SNUP']] = altmax, snup
    h[H['LWALL']] = SBC * c['ewall'] * (tgwall + Ta + 273.15) ** 4 - SBC * c['ewall'] * (Ta + 273.15) ** 4
    h[H['FLUX4']] = SBC * (Ta + 273.15) ** 4
    h[H['LW_AMBIENT']] = (c['ewall'] * SBC * ((Ta + 273.15) ** 4)) / np.pi
    h[H['LW_SUNLIT']] = (c['ewall'] * SBC * ((Ta + tgwall + 273.15) ** 4)) / np.pi
    h[H['LW_REFL']] = (1 - c['ewall']) * 0.5 / np.pi
    h[H['SW_SUNLIT']] = (c['albedo_b'] * (radI * np.cos(alt * DEG)) + (radD * 0.5)) / np.pi
    h[H['SW_SHADED']] = (c['albedo_b'] * radD * 0.5) / np.pi
    ta32, rh32, va32 = (Ta, RH, Ws) if f64 else (np.float32(Ta), np.float32(RH), np.float32(Ws))
    h[H['TA32']] = float(ta32)
    if ta32 <= -999 or rh32 <= -999 or va32 <= -999:
        h[H['UTCI0']:H['UTCI0'] + 7] = -999.0
    else:
        h[H['UTCI0']:H['UTCI0'] + 7] = _utci.coefficients(float(ta32), float(rh32), float(va32))
    h[H['ALBEDO_B']], h[H['ABSK']], h[H['ABSL']] = c['albedo_b'], c['absK'], c['absL']
    h[H['FUP']], h[H['FSIDE']], h[H['FCYL']] = c['Fup'], c['Fside'], c['Fcyl']
    h[H['DIFF_VEG']] = (1.0 - (1.0 - c['trans'])) if f64 else diff_veg(c['trans'])
    h[H['TGWALL']] = tgwall
    h[H['SKY_A']] = SBC * ((Ta + 273.15) ** 4) / np.pi
    h[H['SKY_B']] = (1 - esky) * h[H['SKY_A']]
    pf, pd = patch_rows(alt > 0, esky, Ta, alt, azi, lv, radD, f64=f64) if rows else (None, None)
    if rows:
        bits = np.flatnonzero(pf[:, PF['ELIGIBLE']] != 0)
        for w in range(SW_VIS_WORDS):
            word = bits[(bits >= 32 * w) & (bits < 32 * w + 32)] - 32 * w
            h[H['ELIG'] + w] = float(sum(1 << int(j) for j in word))
    meta = dict(CI=CI, esky=esky, radI=radI, radD=radD, tgwall=tgwall, I0=I0, ea=ea, CI_Tg=citg,
                CI_TgG=citgg, lv=lv, firstdaytime=firstdaytime, timeadd=timeadd)
    return h, pf, pd, off, dz, (CI, firstdaytime, timeadd, timestepdec), meta


def build_day(met, sun, latitude, shape, scale, amaxvalue, *, landcover=0, wall_params=(0.37, -3.41, 15.0),
              constants=CONSTANTS, state=None, psi=None):
    """Tables for the forcing rows of one day (series._run_day, series.py:396-469)."""
    n = met.shape[0]
    dectime = sun['dectime']
    timestepdec = (dectime[1] - dectime[0]) if n > 1 else 1 / 24.
    st = (1., 1, 0., timestepdec) if state is None else tuple(state) + (timestepdec,)
    hours = np.zeros((n, SW_HOUR_STRIDE))
    pf = np.zeros((n, SW_PATCHES, SW_PATCH_F), dtype=np.float32)
    pd = np.zeros((n, SW_PATCH_D, SW_PATCHES))
    ps = np.zeros((n, SW_SEARCH_WORDS), np.uint32)
    offs, dzs, meta = [], [], []
    day_rows, parts = [], []
    start = 0
    for i in range(n):
        if dectime[i] - np.floor(dectime[i]) == 0:
            st = (1.,) + st[1:]
        st_in = st
        h, pf[i], d_i, off, dz, st, m = hour_tables(
            float(sun['altitude'][i]), float(sun['azimuth'][i]), float(sun['zen'][i]), float(sun['jday'][i]),
            float(dectime[i]), float(sun['altmax'][i]), *(float(met[i, k]) for k in (11, 10, 14, 12, 9)),
            met[i, 1], latitude, shape, scale, amaxvalue, st, wall_params=wall_params, constants=constants,
            psi=psi)
        m['state_in'] = st_in
        pd[i] = d_i.T
        if h[H['DAY']]:
            ps[i], part = search_row(pf[i], d_i, _defer=True)
            day_rows.append(i)
            parts.append(part)
        h[H['RAY_START']] = start
        h[H['PATCH_TABLE']] = i
        hours[i] = h
        if len(dz):
            off, dz = pad_steps(off, dz)
            offs.append(off)
            dzs.append(dz)
            start += len(dz)
        meta.append(m)
    if day_rows:
        ps[day_rows] = _with_breakpoints(ps[day_rows], parts)
    ray_off = np.concatenate(offs) if offs else np.zeros((1, 2), np.int32)
    ray_dz = np.concatenate(dzs) if dzs else np.zeros(1)
    t = DayTables(hours, pf, pd, ray_off, ray_dz, meta, ps)
    t._diffuse = diffuse_table(hours, pf)
    t.end_state = st[:3]
    return t
