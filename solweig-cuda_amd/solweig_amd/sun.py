"""Solar position for a forcing table (host, FP64).

Implements the NREL Solar Position Algorithm (Reda & Andreas 2003) with delta_t = 0 and the
fixed 283 K / 1010 mbar refraction term, evaluated the way SOLWEIG 2022a evaluates it:

* ``spa(...)``      - one instant: umep_solweig.py:2529-2678 (julian day 2680-2747, heliocentric
                      position 2750-3069, nutation 3087-3274, obliquity 3277-3298, topocentric
                      terms 3403-3505, range folding 3508-3521).
* ``met_sun(...)``  - per forcing row: Solweig_2015a_metdata_noload (umep_solweig.py:2412-2503):
                      the sun is sampled half a timestep before the stamp, a zenith in (89, 90]
                      is clamped to 89, and ``altmax`` is the 15-minute search from 10:00 local
                      until the altitude stops increasing, redone at each midnight row.

The per-hour arrays it returns drive the ray tables, the Perez sky and the surface-temperature
wave. ``met_sun_device`` evaluates the same per-row algorithm on the device in FP64 (the day
loops' source of solar positions); ``met_sun`` is its host form.
"""
from __future__ import annotations

import calendar
import datetime
import math

import numpy as np

from . import spa_tables as T

_DEG = np.pi / 180.0


def _terms(rows):
    a = np.asarray(rows, dtype=np.float64)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


_L = [_terms(t) for t in (T.L0, T.L1, T.L2, T.L3, T.L4, T.L5)]
_B = [_terms(t) for t in (T.B0, T.B1)]
_R = [_terms(t) for t in (T.R0, T.R1, T.R2, T.R3, T.R4)]
_NUT_ARGS = np.asarray(T.NUTATION_ARGS, dtype=np.float64)
_NUT = np.asarray(T.NUTATION_COEFFS, dtype=np.float64)


def _fold(value, upper):
    value = value - upper * np.floor(value / upper)
    if value < 0:
        value = value + upper
    return value


def _series(groups, jme):
    """sum_k (sum_i A cos(B + C JME)) JME^k / 1e8 with the last group taken un-summed."""
    parts = []
    for k, (a, b, c) in enumerate(groups):
        term = a * np.cos(b + c * jme)
        parts.append(np.sum(term) if k < len(groups) - 1 or len(a) > 1 else term[0])
    total = parts[0]
    for k in range(1, len(parts)):
        total = total + parts[k] * (jme if k == 1 else np.power(jme, k))
    return total / 1e8


def _julian(year, month, day, hour, minute, second, utc):
    if month in (1, 2):
        y, m = year - 1, month + 12
    else:
        y, m = year, month
    ut = ((hour - utc) / 24) + (minute / (60 * 24)) + (second / (60 * 60 * 24))
    d = day + ut
    if year < 1582 or (year == 1582 and (month < 10 or (month == 10 and day <= 4))):
        b = 0
    else:
        a = np.floor(y / 100)
        b = 2 - a + np.floor(a / 4)
    jd = d + b + np.floor(365.25 * (y + 4716)) + np.floor(30.6001 * (m + 1)) - 1524.5
    jc = (jd - 2451545) / 36525
    return jd, jc, jc / 10


def spa(year, month, day, hour, minute, second, utc, latitude, longitude, altitude_m):
    """Topocentric (zenith, azimuth) in degrees for one local instant."""
    jd, jce, jme = _julian(year, month, day, hour, minute, second, utc)
    jc = jce
    # Earth heliocentric longitude, latitude (degrees, folded to [0, 360)) and radius (AU).
    lon = _fold(_series(_L, jme) * 180 / np.pi, 360)
    a0, b0, c0 = _B[0]
    a1, b1, c1 = _B[1]
    lat = (np.sum(a0 * np.cos(b0 + c0 * jme)) + np.sum(a1 * np.cos(b1 + c1 * jme)) * jme) / 1e8
    lat = _fold(lat * 180 / np.pi, 360)
    radius = _series(_R, jme)
    geo_lon = _fold(lon + 180, 360)
    geo_lat = _fold(-lat, 360)
    # Nutation in longitude and obliquity (degrees).
    x0 = (1 / 189474) * np.power(jce, 3) + -0.0019142 * np.power(jce, 2) + 445267.11148 * jce + 297.85036
    x1 = -(1 / 300000) * np.power(jce, 3) + -0.0001603 * np.power(jce, 2) + 35999.05034 * jce + 357.52772
    x2 = (1 / 56250) * np.power(jce, 3) + 0.0086972 * np.power(jce, 2) + 477198.867398 * jce + 134.96298
    x3 = (1 / 327270) * np.power(jce, 3) + -0.0036825 * np.power(jce, 2) + 483202.017538 * jce + 93.27191
    x4 = (1 / 450000) * np.power(jce, 3) + 0.0020708 * np.power(jce, 2) + -1934.136261 * jce + 125.04452
    arg = _NUT_ARGS.dot(np.transpose(np.array([x0, x1, x2, x3, x4]))) * (np.pi / 180)
    dpsi = np.sum((_NUT[:, 0] + _NUT[:, 1] * jce) * np.sin(arg)) / 36000000
    deps = np.sum((_NUT[:, 2] + _NUT[:, 3] * jce) * np.cos(arg)) / 36000000
    u = jme / 10
    poly = (2.45, 5.79, 27.87, 7.12, -39.05, -249.67, -51.38, 1999.25, -1.55, -4680.93, 84381.448)
    eps0 = 0.0
    for k, c in enumerate(poly[:-2]):
        eps0 = eps0 + c * np.power(u, 10 - k)
    eps0 = eps0 + poly[-2] * u + poly[-1]
    eps = eps0 / 3600 + deps
    aberration = -20.4898 / (3600 * radius)
    lam = geo_lon + dpsi + aberration
    nu0 = 280.46061837 + (360.98564736629 * (jd - 2451545)) + (0.000387933 * np.power(jc, 2)) \
        - (np.power(jc, 3) / 38710000)
    nu = _fold(nu0, 360) + dpsi * np.cos(eps * np.pi / 180)
    alpha = np.arctan2(np.sin(lam * np.pi / 180) * np.cos(eps * np.pi / 180)
                       - np.tan(geo_lat * np.pi / 180) * np.sin(eps * np.pi / 180),
                       np.cos(lam * np.pi / 180)) * 180 / np.pi
    alpha = _fold(alpha, 360)
    delta = np.arcsin(np.sin(geo_lat * np.pi / 180) * np.cos(eps * np.pi / 180)
                      + np.cos(geo_lat * np.pi / 180) * np.sin(eps * np.pi / 180)
                      * np.sin(lam * np.pi / 180)) * 180 / np.pi
    h = _fold(nu + longitude - alpha, 360)
    # Topocentric correction (parallax) for an observer at latitude/altitude.
    xi = 8.794 / (3600 * radius)
    uu = np.arctan(0.99664719 * np.tan(latitude * np.pi / 180))
    x = np.cos(uu) + (altitude_m / 6378140) * np.cos(latitude * np.pi / 180)
    y = 0.99664719 * np.sin(uu) + (altitude_m / 6378140) * np.sin(latitude * np.pi / 180)
    dalpha = np.arctan2(-x * np.sin(xi * np.pi / 180) * np.sin(h * np.pi / 180),
                        np.cos(delta * np.pi / 180) - x * np.sin(xi * np.pi / 180) * np.cos(h * np.pi / 180))
    delta_t = np.arctan2((np.sin(delta * np.pi / 180) - y * np.sin(xi * np.pi / 180)) * np.cos(dalpha),
                         np.cos(delta * np.pi / 180) - y * np.sin(xi * np.pi / 180)
                         * np.cos(h * np.pi / 180)) * 180 / np.pi
    h_t = h - dalpha * 180 / np.pi
    e0 = np.arcsin(np.sin(latitude * np.pi / 180) * np.sin(delta_t * np.pi / 180)
                   + np.cos(latitude * np.pi / 180) * np.cos(delta_t * np.pi / 180)
                   * np.cos(h_t * np.pi / 180)) * 180 / np.pi
    refraction = 1.02 / (60 * np.tan((e0 + 10.3 / (e0 + 5.11)) * np.pi / 180))
    zenith = 90 - (e0 + refraction)
    azimuth = np.arctan2(np.sin(h_t * np.pi / 180),
                         np.cos(h_t * np.pi / 180) * np.sin(latitude * np.pi / 180)
                         - np.tan(delta_t * np.pi / 180) * np.cos(latitude * np.pi / 180)) * 180 / np.pi + 180
    return float(zenith), float(_fold(azimuth, 360))


def _fold_v(value, upper):
    value = value - upper * np.floor(value / upper)
    return np.where(value < 0, value + upper, value)


def _series_v(groups, jme):
    """_series for a vector of instants: per instant the same element-wise terms and the same
    contiguous (pairwise) sum of each group."""
    parts = []
    for k, (a, b, c) in enumerate(groups):
        term = a[None, :] * np.cos(b[None, :] + c[None, :] * jme[:, None])
        parts.append(np.sum(term, axis=1) if k < len(groups) - 1 or len(a) > 1 else term[:, 0])
    total = parts[0]
    for k in range(1, len(parts)):
        total = total + parts[k] * (jme if k == 1 else np.power(jme, k))
    return total / 1e8


def spa_batch(stamps, utc, latitude, longitude, altitude_m):
    """``spa`` for a list of datetime instants at one site: (zenith, azimuth) arrays in degrees.
    The arithmetic is spa's, element by element (the nutation arguments per instant)."""
    n = len(stamps)
    jd = np.empty(n)
    for i, st in enumerate(stamps):
        jd[i] = _julian(st.year, st.month, st.day, st.hour, st.minute, 0, utc)[0]
    jc = (jd - 2451545) / 36525
    jce, jme = jc, jc / 10
    lon = _fold_v(_series_v(_L, jme) * 180 / np.pi, 360)
    a0, b0, c0 = _B[0]
    a1, b1, c1 = _B[1]
    lat = (np.sum(a0[None, :] * np.cos(b0[None, :] + c0[None, :] * jme[:, None]), axis=1)
           + np.sum(a1[None, :] * np.cos(b1[None, :] + c1[None, :] * jme[:, None]), axis=1) * jme) / 1e8
    lat = _fold_v(lat * 180 / np.pi, 360)
    radius = _series_v(_R, jme)
    geo_lon = _fold_v(lon + 180, 360)
    geo_lat = _fold_v(-lat, 360)
    x0 = (1 / 189474) * np.power(jce, 3) + -0.0019142 * np.power(jce, 2) + 445267.11148 * jce + 297.85036
    x1 = -(1 / 300000) * np.power(jce, 3) + -0.0001603 * np.power(jce, 2) + 35999.05034 * jce + 357.52772
    x2 = (1 / 56250) * np.power(jce, 3) + 0.0086972 * np.power(jce, 2) + 477198.867398 * jce + 134.96298
    x3 = (1 / 327270) * np.power(jce, 3) + -0.0036825 * np.power(jce, 2) + 483202.017538 * jce + 93.27191
    x4 = (1 / 450000) * np.power(jce, 3) + 0.0020708 * np.power(jce, 2) + -1934.136261 * jce + 125.04452
    X = np.array([x0, x1, x2, x3, x4])
    dpsi = np.empty(n)
    deps = np.empty(n)
    for i in range(n):
        arg = _NUT_ARGS.dot(np.ascontiguousarray(X[:, i])) * (np.pi / 180)
        dpsi[i] = np.sum((_NUT[:, 0] + _NUT[:, 1] * jce[i]) * np.sin(arg)) / 36000000
        deps[i] = np.sum((_NUT[:, 2] + _NUT[:, 3] * jce[i]) * np.cos(arg)) / 36000000
    u = jme / 10
    poly = (2.45, 5.79, 27.87, 7.12, -39.05, -249.67, -51.38, 1999.25, -1.55, -4680.93, 84381.448)
    eps0 = np.zeros(n)
    for k, c in enumerate(poly[:-2]):
        eps0 = eps0 + c * np.power(u, 10 - k)
    eps0 = eps0 + poly[-2] * u + poly[-1]
    eps = eps0 / 3600 + deps
    aberration = -20.4898 / (3600 * radius)
    lam = geo_lon + dpsi + aberration
    nu0 = 280.46061837 + (360.98564736629 * (jd - 2451545)) + (0.000387933 * np.power(jc, 2)) \
        - (np.power(jc, 3) / 38710000)
    nu = _fold_v(nu0, 360) + dpsi * np.cos(eps * np.pi / 180)
    alpha = np.arctan2(np.sin(lam * np.pi / 180) * np.cos(eps * np.pi / 180)
                       - np.tan(geo_lat * np.pi / 180) * np.sin(eps * np.pi / 180),
                       np.cos(lam * np.pi / 180)) * 180 / np.pi
    alpha = _fold_v(alpha, 360)
    delta = np.arcsin(np.sin(geo_lat * np.pi / 180) * np.cos(eps * np.pi / 180)
                      + np.cos(geo_lat * np.pi / 180) * np.sin(eps * np.pi / 180)
                      * np.sin(lam * np.pi / 180)) * 180 / np.pi
    h = _fold_v(nu + longitude - alpha, 360)
    xi = 8.794 / (3600 * radius)
    uu = np.arctan(0.99664719 * np.tan(latitude * np.pi / 180))
    x = np.cos(uu) + (altitude_m / 6378140) * np.cos(latitude * np.pi / 180)
    y = 0.99664719 * np.sin(uu) + (altitude_m / 6378140) * np.sin(latitude * np.pi / 180)
    dalpha = np.arctan2(-x * np.sin(xi * np.pi / 180) * np.sin(h * np.pi / 180),
                        np.cos(delta * np.pi / 180) - x * np.sin(xi * np.pi / 180) * np.cos(h * np.pi / 180))
    delta_t = np.arctan2((np.sin(delta * np.pi / 180) - y * np.sin(xi * np.pi / 180)) * np.cos(dalpha),
                         np.cos(delta * np.pi / 180) - y * np.sin(xi * np.pi / 180)
                         * np.cos(h * np.pi / 180)) * 180 / np.pi
    h_t = h - dalpha * 180 / np.pi
    e0 = np.arcsin(np.sin(latitude * np.pi / 180) * np.sin(delta_t * np.pi / 180)
                   + np.cos(latitude * np.pi / 180) * np.cos(delta_t * np.pi / 180)
                   * np.cos(h_t * np.pi / 180)) * 180 / np.pi
    refraction = 1.02 / (60 * np.tan((e0 + 10.3 / (e0 + 5.11)) * np.pi / 180))
    zenith = 90 - (e0 + refraction)
    azimuth = np.arctan2(np.sin(h_t * np.pi / 180),
                         np.cos(h_t * np.pi / 180) * np.sin(latitude * np.pi / 180)
                         - np.tan(delta_t * np.pi / 180) * np.cos(latitude * np.pi / 180)) * 180 / np.pi + 180
    return zenith, _fold_v(azimuth, 360)


def _instant(stamp, utc, location):
    return spa(stamp.year, stamp.month, stamp.day, stamp.hour, stamp.minute, 0, utc,
               float(location['latitude']), float(location['longitude']), float(location['altitude']))


def _altmax(ymd, utc, location):
    """The day's maximum altitude by the 15-minute search from 10:00 local: the altitude of the
    last sample before it stops increasing (umep_solweig.py:2454-2469), evaluated in blocks."""
    lat, lon, alt = float(location['latitude']), float(location['longitude']), float(location['altitude'])
    fifteen, sunmaximum, zenith = 0., -90., 90.
    while True:
        stamps = []
        f = fifteen
        for _ in range(48):
            f = f + 15. / 1440.
            stamps.append(ymd + datetime.timedelta(days=(60 * 10) / 1440.0 + f))
        zens, _ = spa_batch(stamps, utc, lat, lon, alt)
        for z in zens:
            if not sunmaximum <= 90. - zenith:
                return sunmaximum
            sunmaximum = 90. - zenith
            fifteen = fifteen + 15. / 1440.
            zenith = float(z)
        if not sunmaximum <= 90. - zenith:
            return sunmaximum


def met_sun(met, location, utc):
    """Per-row (altitude deg, azimuth deg, zen rad, jday, dectime, altmax) for a forcing table.

    ``met`` is the [n, >=4] forcing table (year, doy, hour, minute, ...). The rows' instants are
    evaluated together (``spa_batch``, the same values as ``spa`` one at a time).
    """
    met = np.asarray(met, dtype=np.float64)
    n = met.shape[0]
    dectime = met[:, 1] + met[:, 2] / 24 + met[:, 3] / (60 * 24.)
    half = 0 if n == 1 else (dectime[1] - dectime[0]) / 2.
    altmax = np.empty(n)
    jday = np.empty(n)
    stamps = []
    sunmaximum = 0.
    for i in range(n):
        ymd = datetime.datetime(int(met[i, 0]), 1, 1) + datetime.timedelta(int(met[i, 1]) - 1)
        if i == 0 or np.mod(dectime[i], np.floor(dectime[i])) == 0:
            sunmaximum = _altmax(ymd, utc, location)
        altmax[i] = sunmaximum
        stamps.append(ymd + datetime.timedelta(hours=met[i, 2]) + datetime.timedelta(minutes=met[i, 3])
                      - datetime.timedelta(days=half))
        jday[i] = ymd.timetuple().tm_yday
    zenith, azimuth = spa_batch(stamps, utc, float(location['latitude']), float(location['longitude']),
                                float(location['altitude']))
    zenith = np.where((zenith > 89.0) & (zenith <= 90.0), 89.0, zenith)
    altitude = 90. - zenith
    zen = zenith * (np.pi / 180.)
    return dict(altitude=altitude, azimuth=np.asarray(azimuth, dtype=np.float64), zen=zen, jday=jday,
                dectime=dectime, altmax=altmax)


SUN_KEYS = ('altitude', 'azimuth', 'zen', 'jday', 'dectime', 'altmax')


def met_sun_device(met, location, utc, chunk_rows=0, stream=None):
    """``met_sun`` evaluated on the device in FP64 (sw_sun_position), for every row of ``met`` at
    once: the rows are taken in chunks of ``chunk_rows`` (0: one chunk), each chunk as one
    met_sun call. Returns the same dict of host arrays."""
    from . import runtime as rt
    met = np.ascontiguousarray(np.asarray(met, dtype=np.float64))
    n = met.shape[0]
    if n == 0:
        return {k: np.zeros(0) for k in SUN_KEYS}
    dev_met = rt.DeviceArray.from_host(met)
    out = rt.DeviceArray((6, n), np.float64)
    rt.call('sw_sun_position', dev_met.ptr, n, met.shape[1], int(chunk_rows), float(location['latitude']),
            float(location['longitude']), float(location['altitude']), float(utc), out.ptr, rt._s(stream))
    host = out.to_host(stream)
    return {k: host[i].copy() for i, k in enumerate(SUN_KEYS)}


def is_leap(year):
    return calendar.isleap(int(year))


__all__ = ["spa", "met_sun", "met_sun_device", "is_leap", "math"]
