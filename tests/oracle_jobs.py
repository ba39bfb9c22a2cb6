"""Oracle results for the configuration parity tests (CPU only; test infrastructure).

Each job builds a seeded synthetic domain, runs the FP64 oracle on it and returns compact arrays
(mask bytes, Tmrt, UTCI). The GPU tests start these jobs in spawned worker processes so that the
slow NumPy restatement runs while the device work proceeds; nothing here touches the GPU.

Mask byte per pixel and daytime hour, as the shadow pass writes it: bit 0 not building-shadowed
(sh), bit 1 no canopy shadow (vegsh), bit 2 vbshvegsh, bit 3 sunlit wall (wallsun / walls x
ground == 1, umep_solweig.py:152), bit 4 wall facing away from the sun (facesh).
"""
from __future__ import annotations

import datetime

import numpy as np

from oracle import day as oday, umep

# sun positions (azimuth, altitude in degrees) of the long-ray checks, put at hours 5-8 of the day:
# rays of several hundred steps cross the raster, so the skipped windows, both height-bound
# queries and the shadowed-wall threshold drop of the march all take part
LOW_SUNS = ((100.0, 3.0), (250.0, 4.5), (200.0, 3.5), (15.0, 5.0))


def domain(name):
    """The seeded domain of a configuration test."""
    from solweig_amd import synthetic
    if name == 'c0_paved':
        return synthetic.make_c0(landcover=True)
    if name == 'c0_default':
        return synthetic.make_c0()
    if name == 'c1_256':
        return synthetic.make_c1(seed=1, size=256)
    if name == 'c2_2000':
        return synthetic.make_c2(size=2000, days=1)
    if name == 'long_512':
        return synthetic.make_c1(seed=7, size=512)
    if name.startswith('leafoff'):
        # deciduous leaf-off season (psi = 0.5, series.py:416-422) at two latitudes
        date, lat, seed = ((datetime.date(2025, 1, 15), 35.0, 41) if name == 'leafoff_jan'
                           else (datetime.date(2025, 11, 20), 10.0, 42))
        d = synthetic.small_domain(seed=seed, size=64, latitude=lat)
        d.met = synthetic.day_met(date, 1, lat, d.longitude, 0.0,
                                  lambda k: np.random.default_rng(seed * 10 + k), cloudy=True)
        return d
    raise KeyError(name)


def sun(name, d):
    """Per-hour sun of the domain's forcing; the long-ray domain has LOW_SUNS at hours 5-8."""
    from solweig_amd import sun as psun
    s = {k: np.array(v, dtype=np.float64) for k, v in psun.met_sun(d.met, d.location, d.utc).items()}
    if name == 'long_512':
        for h, (az, alt) in zip(range(5, 9), LOW_SUNS):
            s['azimuth'][h], s['altitude'][h], s['zen'][h] = az, alt, (90.0 - alt) * (np.pi / 180.0)
    return s


def mask_byte(shadows, walls, buildings):
    vegsh, sh, vb, _, wallsun, _, facesh, _ = shadows
    with np.errstate(divide='ignore', invalid='ignore'):
        sunwall = (wallsun / walls * buildings) == 1
    return ((sh == 1) * 1 | (vegsh == 1) * 2 | (vb == 1) * 4 | sunwall * 8 | (facesh == 1) * 16).astype(np.uint8)


def _shadow_bytes(q, s, hours):
    out = {}
    for h in hours:
        v = umep.shadows(q['dsm'], q['vegdem'], q['vegdem2'], s['azimuth'][h], s['altitude'][h], 1.0,
                         q['amaxvalue'], q['bush'], q['walls'], q['dirwalls'] * np.pi / 180.)
        out[h] = mask_byte(v, q['walls'], q['buildings'])
    return out


def day_job(name):
    """A full day against the oracle: sky-view masks (packed bits), the mask byte of every daytime
    hour, and Tmrt / UTCI [hours, rows, cols]."""
    d = domain(name)
    s = sun(name, d)
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude, landcover=d.landcover,
                     alb_grid=d.alb_grid, emis_grid=d.emis_grid)
    ref = oday.run_day(p, d.met, s, keep=('Tmrt',))
    day_hours = [h for h in range(d.met.shape[0]) if s['altitude'][h] > 0]
    return dict(svf=tuple(np.packbits(np.asarray(m) != 0, axis=-1) for m in (p['shmat'], p['vegshmat'],
                                                                             p['vbshvegshmat'])),
                masks=_shadow_bytes(p, s, day_hours),
                tmrt=np.stack([r['Tmrt'] for r in ref]), utci=np.stack([r['UTCI'] for r in ref]),
                psi=oday.leaf_psi(d.met[:, 1]))


def shadow_job(name, hours=None):
    """Mask bytes of chosen daytime hours (default: the first hour with the sun at 15-20 degrees)."""
    d = domain(name)
    s = sun(name, d)
    q = oday.shadow_inputs(d.dsm, d.trees, d.dem, d.walls, d.aspect)
    if hours is None:
        hours = [int(np.flatnonzero((s['altitude'] >= 15.0) & (s['altitude'] <= 20.0))[0])]
    return _shadow_bytes(q, s, hours)
