"""Write the golden vectors under tests/golden/ by running the reference's UMEP NumPy module.

Run in the build container only (the reference tree is not present on GPU machines):

    python tests/golden/make_golden.py /root/reference

It imports ``src/umep_solweig.py`` (NumPy only; no CuPy) and evaluates the functions that no
closed form or brute-force check pins on their own:

* ``Solweig_2015a_metdata_noload`` - sun altitude/azimuth/zenith, day of year, dectime, altmax;
* ``clearnessindex_2013b`` and ``Perez_v3`` - per-hour sky scalars;
* ``gvf_2018a`` - the 18-direction ground view factors;
* ``Solweig_2022a_calc`` - whole days of steps, threaded like series.py:396-495;
* the UTCI expression of ``src/kernels/utci.cu`` (translated to Python text and evaluated).

The static per-pixel inputs (walls, aspect, SVF planes and masks) come from this project's
synthetic generator and oracle and are stored beside the outputs, so a test feeds the oracle and
the device path exactly what the reference saw. Only inputs and outputs are written; no reference
source is copied into the fixtures.
"""
from __future__ import annotations

import importlib.util
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, '..', '..'))
sys.path.insert(0, os.path.join(ROOT, 'solweig-cuda_amd'))
sys.path.insert(0, ROOT)

from solweig_amd import synthetic  # noqa: E402
from oracle import day as oday  # noqa: E402


def load_umep(reference):
    spec = importlib.util.spec_from_file_location('umep_solweig', os.path.join(reference, 'src', 'umep_solweig.py'))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def pack_masks(m):
    return np.packbits(np.asarray(m) > 0.5, axis=-1)


def met_cases():
    cases = []
    for lat, lon, doy, utc in ((53.35, -6.26, 172, 0.0), (0.5, 0.0, 80, 0.0), (31.2, 121.5, 355, 8.0),
                               (59.9, 10.7, 150, 1.0), (15.0, -90.0, 300, -6.0)):
        met = np.full((24, 24), -999.0)
        met[:, 0], met[:, 1], met[:, 2], met[:, 3] = 2025, doy, np.arange(24), 0
        cases.append((lat, lon, utc, met))
    return cases


def golden_met(U, out):
    rows = []
    for k, (lat, lon, utc, met) in enumerate(met_cases()):
        loc = {'latitude': lat, 'longitude': lon, 'altitude': 3.0}
        _, alt, azi, zen, jday, _, dectime, altmax = U.Solweig_2015a_metdata_noload(met, loc, utc)
        out[f'met{k}_in'] = met[:, :4]
        out[f'met{k}_loc'] = np.array([lat, lon, utc, 3.0])
        out[f'met{k}_out'] = np.vstack([alt[0], azi[0], zen[0], jday[0], dectime, altmax[0]])
        rows.append(k)
    out['met_cases'] = np.array(rows)


def golden_sky(U, out):
    rng = np.random.default_rng(11)
    ci, pz = [], []
    for lat in (5.0, 15.0, 25.0, 35.0, 45.0, 53.35, 65.0, 75.0, 85.0):
        for jday in (20, 100, 200, 300, 350):
            zen = rng.uniform(0.1, 1.5)
            Ta, RH, radG, P = rng.uniform(-5, 35), rng.uniform(0.2, 0.95), rng.uniform(20, 900), rng.choice([-999.0, 101.3])
            I0, CI, Kt, I0et, CIu = U.clearnessindex_2013b(zen, jday, Ta, RH, radG, {'latitude': lat}, P)
            ci.append([zen, jday, Ta, RH, radG, lat, P, I0, CI, Kt, I0et, CIu])
    out['clearness'] = np.array(ci, dtype=np.float64)
    for k in range(40):
        zen = rng.uniform(5, 85)
        azi = rng.uniform(0, 360)
        radD = rng.uniform(5, 300)
        radI = rng.uniform(0, 900) * (k % 5 != 0)
        jday = int(rng.integers(1, 366))
        lv, clear, bright = U.Perez_v3(zen, azi, radD, radI, jday, 1, 2)
        pz.append(np.concatenate([[zen, azi, radD, radI, jday, clear, bright], lv[:, 2]]))
        if k == 0:
            out['perez_angles'] = lv[:, :2]
    out['perez'] = np.array(pz)


def static_fixture(prefix, d, out):
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude, landcover=d.landcover,
                     alb_grid=d.alb_grid, emis_grid=d.emis_grid)
    out[prefix + 'dsm'], out[prefix + 'trees'], out[prefix + 'dem'] = d.dsm, d.trees, d.dem
    out[prefix + 'walls'], out[prefix + 'aspect'] = d.walls, d.aspect
    if d.landcover is not None:
        out[prefix + 'landcover'] = d.landcover
    if d.alb_grid is not None:
        out[prefix + 'alb_grid'], out[prefix + 'emis_grid'] = d.alb_grid, d.emis_grid
    out[prefix + 'loc'] = np.array([d.latitude, d.longitude, d.utc, d.site_altitude])
    out[prefix + 'met'] = d.met
    out[prefix + 'shmat'] = pack_masks(p['shmat'])
    out[prefix + 'vegmat'] = pack_masks(p['vegshmat'])
    out[prefix + 'vbmat'] = pack_masks(p['vbshvegshmat'])
    out[prefix + 'svf15'] = np.stack([p[k] for k in SVF_KEYS]).astype(np.float32)
    return p


SVF_KEYS = ('svf', 'svfaveg', 'svfE', 'svfEaveg', 'svfEveg', 'svfN', 'svfNaveg', 'svfNveg',
            'svfS', 'svfSaveg', 'svfSveg', 'svfveg', 'svfW', 'svfWaveg', 'svfWveg')
SAMPLE_HOURS = (0, 4, 6, 9, 12, 15, 18, 20)
STEP_KEYS = ('Tmrt', 'Kdown', 'Kup', 'Ldown', 'Lup', 'Tg', 'shadow', 'Keast', 'Ksouth', 'Kwest',
             'Knorth', 'Least', 'Lsouth', 'Lwest', 'Lnorth', 'KsideI', 'TgOut', 'Lside', 'KsideD',
             'dRad', 'Kside')


def golden_day(U, prefix, d, out, gvf_hours=(8, 14)):
    p = static_fixture(prefix, d, out)
    loc = d.location
    met = d.met
    _, alt, azi, zen, jday, _, dectime, altmax = U.Solweig_2015a_metdata_noload(met, loc, d.utc)
    rows, cols = d.shape
    psi = oday.leaf_psi(met[:, 1])
    z = np.zeros((rows, cols))
    state = dict(Tgmap1=z, Tgmap1E=z, Tgmap1S=z, Tgmap1W=z, Tgmap1N=z, TgOut1=z)
    CI, firstdaytime, timeadd = 1., 1., 0.
    timestepdec = dectime[1] - dectime[0]
    Twater = float(np.mean(met[:, 11])) if p['landcover'] else None
    recs = {k: [] for k in STEP_KEYS}
    scal = []
    for i in range(met.shape[0]):
        if dectime[i] - np.floor(dectime[i]) == 0:
            CI = 1.
        args = dict(i=i, dsm=p['dsm'], scale=1.0, rows=rows, cols=cols,
                    svf=p['svf'], svfN=p['svfN'], svfW=p['svfW'], svfE=p['svfE'], svfS=p['svfS'],
                    svfveg=p['svfveg'], svfNveg=p['svfNveg'], svfEveg=p['svfEveg'], svfSveg=p['svfSveg'],
                    svfWveg=p['svfWveg'], svfaveg=p['svfaveg'], svfEaveg=p['svfEaveg'],
                    svfSaveg=p['svfSaveg'], svfWaveg=p['svfWaveg'], svfNaveg=p['svfNaveg'],
                    vegdem=p['vegdem'], vegdem2=p['vegdem2'], albedo_b=0.2, absK=0.7, absL=0.95,
                    ewall=0.9, Fside=0.22, Fup=0.06, Fcyl=0.28, altitude=float(alt[0, i]),
                    azimuth=float(azi[0, i]), zen=float(zen[0, i]), jday=float(jday[0, i]),
                    usevegdem=1, onlyglobal=1, buildings=p['buildings'], location=loc,
                    psi=float(psi[i]), landcover=p['landcover'], lc_grid=p['lc_grid'],
                    dectime=float(dectime[i]), altmax=float(altmax[0, i]), dirwalls=p['dirwalls'],
                    walls=p['walls'], cyl=1, elvis=0, Ta=float(met[i, 11]), RH=float(met[i, 10]),
                    radG=float(met[i, 14]), radD=float(met[i, 21]), radI=float(met[i, 22]),
                    P=float(met[i, 12]), amaxvalue=p['amaxvalue'], bush=p['bush'], Twater=Twater,
                    TgK=p['TgK'], Tstart=p['Tstart'], alb_grid=p['alb_grid'], emis_grid=p['emis_grid'],
                    TgK_wall=p['TgK_wall'], Tstart_wall=p['Tstart_wall'], TmaxLST=p['TmaxLST'],
                    TmaxLST_wall=p['TmaxLST_wall'], first=1.0, second=22.0, svfalfa=p['svfalfa'],
                    svfbuveg=p['svfbuveg'], firstdaytime=firstdaytime, timeadd=timeadd,
                    timestepdec=timestepdec, CI=CI, diffsh=p['diffsh'], shmat=p['shmat'],
                    vegshmat=p['vegshmat'], vbshvegshmat=p['vbshvegshmat'], anisotropic_sky=1,
                    asvf=p['asvf'], patch_option=2, **state)
        r = U.Solweig_2022a_calc(**args)
        (Tmrt, Kdown, Kup, Ldown, Lup, Tg, ea, esky, I0, CI, shadow, firstdaytime, timestepdec,
         timeadd, Tgmap1, Tgmap1E, Tgmap1S, Tgmap1W, Tgmap1N, Keast, Ksouth, Kwest, Knorth, Least,
         Lsouth, Lwest, Lnorth, KsideI, TgOut1, TgOut, radI, radD, Lside, L_patches, CI_Tg, CI_TgG,
         KsideD, dRad, Kside) = r
        state = dict(Tgmap1=Tgmap1, Tgmap1E=Tgmap1E, Tgmap1S=Tgmap1S, Tgmap1W=Tgmap1W,
                     Tgmap1N=Tgmap1N, TgOut1=TgOut1)
        vals = dict(Tmrt=Tmrt, Kdown=Kdown, Kup=Kup, Ldown=Ldown, Lup=Lup, Tg=Tg, shadow=shadow,
                    Keast=Keast, Ksouth=Ksouth, Kwest=Kwest, Knorth=Knorth, Least=Least,
                    Lsouth=Lsouth, Lwest=Lwest, Lnorth=Lnorth, KsideI=KsideI, TgOut=TgOut,
                    Lside=Lside, KsideD=KsideD, dRad=dRad, Kside=Kside)
        for k in STEP_KEYS:
            recs[k].append(np.broadcast_to(np.asarray(vals[k], dtype=np.float64), (rows, cols)))
        scal.append([float(np.asarray(v).item()) if np.ndim(v) == 0 or np.size(v) == 1 else np.nan
                     for v in (ea, esky, I0, CI, firstdaytime, timeadd, radI, radD, CI_Tg, CI_TgG)])
        if i in gvf_hours and alt[0, i] > 0:
            sh_out = U.shadowingfunction_wallheight_23(p['dsm'], p['vegdem'], p['vegdem2'], float(azi[0, i]),
                                                       float(alt[0, i]), 1.0, p['amaxvalue'], p['bush'],
                                                       p['walls'], p['dirwalls'] * np.pi / 180.)
            out[prefix + f'shadow{i}'] = np.stack(sh_out).astype(np.float32)
            sh_ = sh_out[1] - (1 - sh_out[0]) * (1 - psi[i])
            tg = np.asarray(Tg, dtype=np.float64) / 1.0
            g = U.gvf_2018a(sh_out[4], p['walls'], p['buildings'], 1.0, sh_, 1.0, 22.0, p['dirwalls'],
                            np.array(tg, copy=True), 3.5, float(met[i, 11]), p['emis_grid'], 0.9,
                            p['alb_grid'], 5.67051e-8, 0.2, rows, cols, Twater, p['lc_grid'], p['landcover'])
            out[prefix + f'gvf{i}'] = np.stack(g).astype(np.float32)
            out[prefix + f'gvf{i}_Tg'] = tg.astype(np.float32)
            out[prefix + f'gvf{i}_shadow'] = np.asarray(sh_, dtype=np.float32)
    keep = np.array(SAMPLE_HOURS)
    out[prefix + 'step_hours'] = keep
    for k in STEP_KEYS:
        series = np.stack(recs[k])
        out[prefix + 'step_' + k] = (series if k == 'Tmrt' else series[keep]).astype(np.float32)
    out[prefix + 'step_scalars'] = np.array(scal)
    out[prefix + 'sun'] = np.vstack([alt[0], azi[0], zen[0], jday[0], dectime, altmax[0]])


def golden_utci(reference, out):
    src = open(os.path.join(reference, 'src', 'kernels', 'utci.cu')).read()
    body = src[src.index('float tk = Ta + 273.15f;'):src.index('output[i] = out;')]
    py = []
    for stmt in body.split(';'):
        stmt = ' '.join(stmt.split())
        if not stmt or stmt.startswith('}') or '=' not in stmt:
            continue
        stmt = re.sub(r'^float\s+', '', stmt)
        stmt = re.sub(r'powf\(([^,]+),\s*([-0-9.]+)f\)', r'(\1)**(\2)', stmt)
        stmt = stmt.replace('logf(', 'np.log(').replace('expf(', 'np.exp(')
        stmt = re.sub(r'([0-9.]+(?:E[-+]?[0-9]+)?)f\b', r'\1', stmt)
        m = re.match(r'^(\w+)\s*(\+?=)\s*(.*)$', stmt)
        py.append(f'{m.group(1)} {m.group(2)} ({m.group(3)})')
    code = '\n'.join(py)
    rng = np.random.default_rng(5)
    n = 400
    Ta = rng.uniform(-20, 45, n)
    RH = rng.uniform(5, 100, n)
    Tmrt = Ta + rng.uniform(-20, 60, n)
    va = rng.uniform(0.5, 15, n)
    res = np.empty(n)
    for k in range(n):
        env = {'np': np, 'Ta': Ta[k], 'RH': RH[k], 'Tmrt': Tmrt[k], 'va10m': va[k]}
        exec(code, env)
        res[k] = env['out']
    out['utci'] = np.vstack([Ta, RH, Tmrt, va, res])


def main():
    reference = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
    U = load_umep(reference)
    out = {}
    golden_met(U, out)
    golden_sky(U, out)
    golden_utci(reference, out)
    d = synthetic.small_domain(seed=21, size=32)
    golden_day(U, 'dayA_', d, out)
    d2 = synthetic.small_domain(seed=22, size=24, trees=False, landcover=True)
    golden_day(U, 'dayB_', d2, out, gvf_hours=(10,))
    d3 = synthetic.small_domain(seed=23, size=24, maps=True)
    golden_day(U, 'dayC_', d3, out, gvf_hours=(12,))
    np.savez_compressed(os.path.join(HERE, 'umep_golden.npz'), **out)
    print('wrote', os.path.join(HERE, 'umep_golden.npz'), sum(v.nbytes for v in out.values()), 'bytes raw')


if __name__ == '__main__':
    main()
