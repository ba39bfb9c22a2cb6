"""Solar position restated in the stepwise form of umep_solweig.py (oracle; test only).

Each helper follows one stage of the original (line ranges cited); the periodic-term tables are
the published NREL SPA values (Reda & Andreas 2003), loaded from the package's table module.
"""
from __future__ import annotations

import datetime
import importlib.util
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_TABLES = os.path.join(_HERE, '..', 'solweig-cuda_amd', 'solweig_amd', 'spa_tables.py')
_spec = importlib.util.spec_from_file_location('_oracle_spa_tables', _TABLES)
T = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(T)


def _rng(var, hi):
    """set_to_range (umep_solweig.py:3508-3521)."""
    var = var - hi * np.floor(var / hi)
    if var < 0:
        var = var + hi
    return var


def julian(t):
    """umep_solweig.py:2680-2747 (Gregorian branch, delta_t = 0)."""
    if t['month'] in (1, 2):
        Y, M = t['year'] - 1, t['month'] + 12
    else:
        Y, M = t['year'], t['month']
    ut = ((t['hour'] - t['UTC']) / 24) + (t['min'] / (60 * 24)) + (t['sec'] / (60 * 60 * 24))
    D = t['day'] + ut
    A = np.floor(Y / 100)
    B = 2 - A + np.floor(A / 4)
    j = {'day': D + B + np.floor(365.25 * (Y + 4716)) + np.floor(30.6001 * (M + 1)) - 1524.5}
    j['ephemeris_day'] = j['day'] + 0 / 86400
    j['century'] = (j['day'] - 2451545) / 36525
    j['ephemeris_century'] = (j['ephemeris_day'] - 2451545) / 36525
    j['ephemeris_millenium'] = j['ephemeris_century'] / 10
    return j


def _sum(rows, jme):
    a = np.asarray(rows, dtype=np.float64)
    return np.sum(a[:, 0] * np.cos(a[:, 1] + (a[:, 2] * jme)))


def heliocentric(j):
    """umep_solweig.py:2750-3069."""
    JME = j['ephemeris_millenium']
    L = [_sum(t, JME) for t in (T.L0, T.L1, T.L2, T.L3, T.L4)]
    L5 = np.asarray(T.L5)[0]
    L5 = L5[0] * np.cos(L5[1] + L5[2] * JME)
    lon = (L[0] + (L[1] * JME) + (L[2] * np.power(JME, 2)) + (L[3] * np.power(JME, 3))
           + (L[4] * np.power(JME, 4)) + (L5 * np.power(JME, 5))) / 1e8
    lon = _rng(lon * 180 / np.pi, 360)
    lat = (_sum(T.B0, JME) + (_sum(T.B1, JME) * JME)) / 1e8
    lat = _rng(lat * 180 / np.pi, 360)
    R = [_sum(t, JME) for t in (T.R0, T.R1, T.R2, T.R3)]
    R4 = np.asarray(T.R4)[0]
    R4 = R4[0] * np.cos(R4[1] + R4[2] * JME)
    rad = (R[0] + (R[1] * JME) + (R[2] * np.power(JME, 2)) + (R[3] * np.power(JME, 3))
           + (R4 * np.power(JME, 4))) / 1e8
    return {'longitude': lon, 'latitude': lat, 'radius': rad}


def nutation(j):
    """umep_solweig.py:3087-3274."""
    JCE = j['ephemeris_century']
    polys = ((1 / 189474, -0.0019142, 445267.11148, 297.85036),
             (-(1 / 300000), -0.0001603, 35999.05034, 357.52772),
             (1 / 56250, 0.0086972, 477198.867398, 134.96298),
             (1 / 327270, -0.0036825, 483202.017538, 93.27191),
             (1 / 450000, 0.0020708, -1934.136261, 125.04452))
    X = np.array([p[0] * np.power(JCE, 3) + p[1] * np.power(JCE, 2) + p[2] * JCE + p[3] for p in polys])
    Y = np.asarray(T.NUTATION_ARGS, dtype=np.float64)
    C = np.asarray(T.NUTATION_COEFFS, dtype=np.float64)
    arg = Y.dot(np.transpose(X)) * (np.pi / 180)
    return {'longitude': np.sum((C[:, 0] + (C[:, 1] * JCE)) * np.sin(arg)) / 36000000,
            'obliquity': np.sum((C[:, 2] + (C[:, 3] * JCE)) * np.cos(arg)) / 36000000}


def sun_position(t, loc):
    """Zenith and azimuth in degrees (umep_solweig.py:2529-2678, 3277-3505)."""
    j = julian(t)
    e = heliocentric(j)
    g_lon = _rng(e['longitude'] + 180, 360)
    g_lat = _rng(-e['latitude'], 360)
    nu = nutation(j)
    p = (2.45, 5.79, 27.87, 7.12, -39.05, -249.67, -51.38, 1999.25, -1.55, -4680.93, 84381.448)
    U = j['ephemeris_millenium'] / 10
    mean_obl = p[0] * np.power(U, 10) + p[1] * np.power(U, 9) + p[2] * np.power(U, 8) + \
        p[3] * np.power(U, 7) + p[4] * np.power(U, 6) + p[5] * np.power(U, 5) + \
        p[6] * np.power(U, 4) + p[7] * np.power(U, 3) + p[8] * np.power(U, 2) + p[9] * U + p[10]
    obl = (mean_obl / 3600) + nu['obliquity']
    aberr = -20.4898 / (3600 * e['radius'])
    lam = g_lon + nu['longitude'] + aberr
    JD, JC = j['day'], j['century']
    st = 280.46061837 + (360.98564736629 * (JD - 2451545)) + (0.000387933 * np.power(JC, 2)) \
        - (np.power(JC, 3) / 38710000)
    st = _rng(st, 360) + (nu['longitude'] * np.cos(obl * np.pi / 180))
    ra = np.arctan2((np.sin(lam * np.pi / 180) * np.cos(obl * np.pi / 180))
                    - (np.tan(g_lat * np.pi / 180) * np.sin(obl * np.pi / 180)),
                    np.cos(lam * np.pi / 180)) * 180 / np.pi
    ra = _rng(ra, 360)
    dec = np.arcsin((np.sin(g_lat * np.pi / 180) * np.cos(obl * np.pi / 180))
                    + (np.cos(g_lat * np.pi / 180) * np.sin(obl * np.pi / 180)
                       * np.sin(lam * np.pi / 180))) * 180 / np.pi
    H = _rng(st + loc['longitude'] - ra, 360)
    xi = 8.794 / (3600 * e['radius'])
    u = np.arctan(0.99664719 * np.tan(loc['latitude'] * np.pi / 180))
    x = np.cos(u) + ((loc['altitude'] / 6378140) * np.cos(loc['latitude'] * np.pi / 180))
    y = (0.99664719 * np.sin(u)) + ((loc['altitude'] / 6378140) * np.sin(loc['latitude'] * np.pi / 180))
    dra = np.arctan2(-x * np.sin(xi * np.pi / 180) * np.sin(H * np.pi / 180),
                     np.cos(dec * np.pi / 180) - (x * np.sin(xi * np.pi / 180) * np.cos(H * np.pi / 180)))
    tdec = np.arctan2((np.sin(dec * np.pi / 180) - (y * np.sin(xi * np.pi / 180))) * np.cos(dra),
                      np.cos(dec * np.pi / 180) - (y * np.sin(xi * np.pi / 180)) * np.cos(H * np.pi / 180)) \
        * 180 / np.pi
    tH = H - dra * 180 / np.pi
    arg = (np.sin(loc['latitude'] * np.pi / 180) * np.sin(tdec * np.pi / 180)) + \
        (np.cos(loc['latitude'] * np.pi / 180) * np.cos(tdec * np.pi / 180) * np.cos(tH * np.pi / 180))
    e0 = np.arcsin(arg) * 180 / np.pi
    refr = 1.02 / (60 * np.tan((e0 + (10.3 / (e0 + 5.11))) * np.pi / 180))
    zen = 90 - (e0 + refr)
    az = (np.arctan2(np.sin(tH * np.pi / 180),
                     (np.cos(tH * np.pi / 180) * np.sin(loc['latitude'] * np.pi / 180))
                     - (np.tan(tdec * np.pi / 180) * np.cos(loc['latitude'] * np.pi / 180)))
          * 180 / np.pi) + 180
    return {'zenith': float(np.asarray(zen).item()), 'azimuth': float(np.asarray(_rng(az, 360)).item())}


def metdata(met, loc, utc):
    """Solweig_2015a_metdata_noload (umep_solweig.py:2412-2503)."""
    n = met.shape[0]
    dectime = met[:, 1] + met[:, 2] / 24 + met[:, 3] / (60 * 24.)
    half = 0 if n == 1 else (dectime[1] - dectime[0]) / 2.
    t = {'sec': 0, 'UTC': utc}
    alt, azi, zen, jday, altmax = (np.empty(n) for _ in range(5))
    smax = 0.
    for i in range(n):
        ymd = datetime.datetime(int(met[i, 0]), 1, 1) + datetime.timedelta(int(met[i, 1]) - 1)
        if i == 0 or np.mod(dectime[i], np.floor(dectime[i])) == 0:
            fifteen, smax, sm = 0., -90., {'zenith': 90.}
            while smax <= 90. - sm['zenith']:
                smax = 90. - sm['zenith']
                fifteen = fifteen + 15. / 1440.
                s = ymd + datetime.timedelta(days=(60 * 10) / 1440.0 + fifteen)
                t.update(year=s.year, month=s.month, day=s.day, hour=s.hour, min=s.minute)
                sm = sun_position(t, loc)
        altmax[i] = smax
        s = ymd + datetime.timedelta(hours=met[i, 2]) + datetime.timedelta(minutes=met[i, 3]) \
            - datetime.timedelta(days=half)
        t.update(year=s.year, month=s.month, day=s.day, hour=s.hour, min=s.minute)
        sun = sun_position(t, loc)
        if 89.0 < sun['zenith'] <= 90.0:
            sun['zenith'] = 89.0
        alt[i] = 90. - sun['zenith']
        azi[i] = sun['azimuth']
        zen[i] = sun['zenith'] * (np.pi / 180.)
        jday[i] = ymd.timetuple().tm_yday
    return dict(altitude=alt, azimuth=azi, zen=zen, jday=jday, dectime=dectime, altmax=altmax)
