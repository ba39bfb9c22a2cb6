"""Static preparation and the day loop around the oracle step (oracle; test infrastructure only).

``prepare`` restates the static derivations of series.py:252-349 (DSM/DEM raise, vegetation
DSMs and bush, amaxvalue, ground mask, land-cover or default maps, SVF and derived planes);
``run_day`` restates series.py:396-495 for one day of forcing rows.
"""
from __future__ import annotations

import numpy as np

from . import svf as osvf
from . import umep
from . import utci as outci

LANDCOVER_TABLE = np.array([  # code, albedo, emissivity, TgK, Tstart, TmaxLST (landcoverclasses_2016a.txt)
    [2, 0.18, 0.95, 0.58, -9.78, 15.0],
    [1, 0.18, 0.95, 0.58, -9.78, 15.0],
    [0, 0.20, 0.95, 0.37, -3.41, 15.0],
    [7, 0.05, 0.98, 0.00, 0.00, 12.0],
    [5, 0.16, 0.94, 0.21, -3.38, 14.0],
    [6, 0.25, 0.94, 0.33, -3.01, 14.0],
    [99, 0.20, 0.90, 0.58, -3.41, 15.0]])

CONSTANTS = dict(albedo_b=0.2, ewall=0.9, absK=0.7, absL=0.95, Fside=0.22, Fup=0.06, Fcyl=0.28,
                 elvis=0, trans=0.03, first=1.0, second=22.0)


def normalise_landcover(values):
    """series.py:505-513."""
    src = np.asarray(values, dtype=float)
    r = np.rint(src)
    bad = (~np.isfinite(src)) | (~np.isclose(src, r)) | (r < 1) | (r > 7)
    r[bad] = 6
    r[(r == 3) | (r == 4)] = 5
    return r.astype(np.float32)


def tg_maps(lc):
    """Tgmaps_v1 (umep_solweig.py:1603-1628) on a normalised grid."""
    out = {k: np.copy(lc).astype(np.float64) for k in ('TgK', 'Tstart', 'alb', 'emis', 'TmaxLST')}
    for code in np.unique(lc):
        row = LANDCOVER_TABLE[LANDCOVER_TABLE[:, 0] == code][0]
        sel = lc == code
        out['alb'][sel], out['emis'][sel] = row[1], row[2]
        out['TgK'][sel], out['Tstart'][sel], out['TmaxLST'][sel] = row[3], row[4], row[5]
    wall = LANDCOVER_TABLE[LANDCOVER_TABLE[:, 0] == 99][0]
    return out, (wall[3], wall[4], wall[5])


def shadow_inputs(dsm, trees, dem, walls, aspect):
    """The static inputs of the sun-ray march only (series.py:252-292): raised DSM and DEM,
    vegetation DSMs, bush, amaxvalue, ground mask, walls and aspect (degrees)."""
    a = np.asarray(dsm, dtype=np.float64)
    t = np.asarray(trees, dtype=np.float64).copy()
    g = np.asarray(dem, dtype=np.float64)
    dsm_raise = max(0.0, -float(a.min()))
    dem_raise = max(0.0, -float(g.min()))
    if dsm_raise:
        a = a + dsm_raise
    if dem_raise:
        g = g + dem_raise
    t[t < 0.] = 0.
    vegdsm = t + a
    vegdsm[vegdsm == a] = 0
    vegdsm2 = t * 0.25 + a
    vegdsm2[vegdsm2 == a] = 0
    bush = np.logical_not(vegdsm2 * vegdsm) * vegdsm
    amaxvalue = max(a.max() - a.min(), t.max())
    buildings = a - g
    buildings[buildings < 2.] = 1.
    buildings[buildings >= 2.] = 0.
    return dict(dsm=a, dem=g, trees=t, vegdem=vegdsm, vegdem2=vegdsm2, bush=bush, amaxvalue=float(amaxvalue),
                buildings=buildings, walls=np.asarray(walls, dtype=np.float64),
                dirwalls=np.asarray(aspect, dtype=np.float64))


def prepare(dsm, trees, dem, walls, aspect, latitude, *, scale=1.0, landcover=None, alb_grid=None,
            emis_grid=None, svf=None):
    """Static inputs of the step (series.py:252-365). ``svf`` may carry precomputed
    (15-tuple + vegmat, vbmat, shmat, total) from ``oracle.svf.svf_calculator``."""
    q = shadow_inputs(dsm, trees, dem, walls, aspect)
    a, t, g = q['dsm'], q['trees'], q['dem']
    rows, cols = a.shape
    p = dict(CONSTANTS)
    z = np.zeros((rows, cols), dtype=np.float32)
    if landcover is not None:
        lc = normalise_landcover(landcover)
        maps, (tgk_w, ts_w, tmax_w) = tg_maps(lc)
        p.update(landcover=1, lc_grid=lc.astype(np.float64),
                 TgK=maps['TgK'].astype(np.float32).astype(np.float64),
                 Tstart=maps['Tstart'].astype(np.float32).astype(np.float64),
                 alb_grid=maps['alb'].astype(np.float32).astype(np.float64),
                 emis_grid=maps['emis'].astype(np.float32).astype(np.float64),
                 TmaxLST=maps['TmaxLST'].astype(np.float32).astype(np.float64),
                 TgK_wall=float(np.float32(tgk_w)), Tstart_wall=float(np.float32(ts_w)),
                 TmaxLST_wall=float(np.float32(tmax_w)))
    else:
        f32 = lambda v: (z + np.float32(v)).astype(np.float64)
        p.update(landcover=0, lc_grid=np.zeros((rows, cols)), TgK=f32(0.37), Tstart=f32(-3.41),
                 alb_grid=f32(0.15), emis_grid=f32(0.95), TmaxLST=f32(15.), TgK_wall=0.37,
                 Tstart_wall=-3.41, TmaxLST_wall=15.)
    if alb_grid is not None:
        p['alb_grid'] = np.asarray(alb_grid, dtype=np.float32).astype(np.float64)
    if emis_grid is not None:
        p['emis_grid'] = np.asarray(emis_grid, dtype=np.float32).astype(np.float64)
    vegdsm, vegdsm2, bush, amaxvalue, buildings = (q[k] for k in ('vegdem', 'vegdem2', 'bush', 'amaxvalue',
                                                                   'buildings'))
    if svf is None:
        svf = osvf.svf_calculator(amaxvalue, a, vegdsm, vegdsm2, bush, scale, p['trans'])
    planes = tuple(np.asarray(v).astype(np.float32).astype(np.float64) for v in svf[:15])
    (s, saveg, sE, sEaveg, sEveg, sN, sNaveg, sNveg, sS, sSaveg, sSveg, sveg, sW, sWaveg, sWveg) = planes
    vegmat, vbmat, shmat = svf[15], svf[16], svf[17]
    svfbuveg = s - (1.0 - sveg) * (1.0 - p['trans'])
    asvf = np.arccos(np.sqrt(s))
    tmp = s + sveg - 1.0
    tmp[tmp < 0.0] = 0.0
    with np.errstate(divide='ignore'):
        svfalfa = np.arcsin(np.exp(np.log(1.0 - tmp) / 2.0))
    p.update(dsm=a, dem=g, vegdem=vegdsm, vegdem2=vegdsm2, bush=bush, amaxvalue=float(amaxvalue),
             buildings=buildings, walls=q['walls'], dirwalls=q['dirwalls'], scale=float(scale),
             latitude=float(latitude),
             svf=s, svfaveg=saveg, svfE=sE, svfEaveg=sEaveg, svfEveg=sEveg, svfN=sN, svfNaveg=sNaveg,
             svfNveg=sNveg, svfS=sS, svfSaveg=sSaveg, svfSveg=sSveg, svfveg=sveg, svfW=sW,
             svfWaveg=sWaveg, svfWveg=sWveg, svfbuveg=svfbuveg, asvf=asvf, svfalfa=svfalfa,
             shmat=shmat, vegshmat=vegmat, vbshvegshmat=vbmat,
             diffsh=osvf.diffsh(shmat, vegmat, p['trans']))
    return p


def with_transmissivity(p, trans):
    """p with canopy transmissivity ``trans`` in svfbuveg and diffsh (series.py:343-349)."""
    q = dict(p)
    q['trans'] = float(trans)
    q['svfbuveg'] = p['svf'] - (1.0 - p['svfveg']) * (1.0 - q['trans'])
    q['diffsh'] = osvf.diffsh(p['shmat'], p['vegshmat'], q['trans'])
    return q


def leaf_psi(doy):
    """series.py:416-422 (deciduous, leaf-on strictly between DOY 97 and 300)."""
    return np.where((doy > 97) & (doy < 300), 3. / 100., .5)


def run_day(p, met, sun, *, keep=('Tmrt',), psi=None):
    """One day of hourly steps (series.py:396-495). ``sun`` holds per-row altitude, azimuth,
    zen, jday, dectime, altmax. Returns a list of per-hour dicts with ``keep`` outputs + UTCI.
    ``psi`` overrides the seasonal canopy transmissivity of series.py:416-422."""
    rows, cols = p['dsm'].shape
    n = met.shape[0]
    psi = leaf_psi(met[:, 1]) if psi is None else np.full(n, float(psi))
    dectime = sun['dectime']
    timestepdec = (dectime[1] - dectime[0]) if n > 1 else 1 / 24.
    zero = np.zeros((rows, cols))
    state = dict(CI=1., firstdaytime=1., timeadd=0., timestepdec=timestepdec, Tgmap1=zero,
                 Tgmap1E=zero, Tgmap1S=zero, Tgmap1W=zero, Tgmap1N=zero, TgOut1=zero)
    if p['landcover']:
        p = dict(p, Twater=float(np.mean(met[:, 11])))
    ground = p['buildings'] == 1
    out = []
    for i in range(n):
        if dectime[i] - np.floor(dectime[i]) == 0:
            state['CI'] = 1.
        hour = dict(altitude=float(sun['altitude'][i]), azimuth=float(sun['azimuth'][i]),
                    zen=float(sun['zen'][i]), jday=float(sun['jday'][i]), dectime=float(dectime[i]),
                    altmax=float(sun['altmax'][i]), Ta=float(met[i, 11]), RH=float(met[i, 10]),
                    radG=float(met[i, 14]), radD=float(met[i, 21]), radI=float(met[i, 22]),
                    P=float(met[i, 12]), psi=float(psi[i]))
        res, state = umep.step(p, hour, state)
        u = outci.utci(np.float32(met[i, 11]), np.float32(met[i, 10]), res['Tmrt'], np.float32(met[i, 9]))
        u = np.where(ground, u, np.nan)
        rec = {k: res[k] for k in keep if k in res}
        rec['UTCI'] = u
        out.append(rec)
    return out
