"""The per-step drop-in (model.Solweig_2022a_calc) against UMEP's own step (golden vectors, GPU)."""
import numpy as np
import pytest

from golden_inputs import static_from_golden, sun_from_golden

pytestmark = pytest.mark.gpu

STEP_KEYS = ('Tmrt', 'Kdown', 'Kup', 'Ldown', 'Lup', 'Tg', 'shadow', 'Keast', 'Ksouth', 'Kwest',
             'Knorth', 'Least', 'Lsouth', 'Lwest', 'Lnorth', 'KsideI', 'TgOut', 'Lside', 'KsideD',
             'dRad', 'Kside')


@pytest.mark.parametrize('prefix', ['dayA_', 'dayB_', 'dayC_'])
def test_step_api_matches_umep(golden, prefix):
    from solweig_amd import model, runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    p = static_from_golden(golden, prefix)
    met = golden[prefix + 'met']
    s = sun_from_golden(golden, prefix)
    loc = golden[prefix + 'loc']
    rows, cols = p['dsm'].shape
    print(f'step API vs UMEP {prefix}', flush=True)
    args = dict(p)
    args.update(rows=rows, cols=cols, location={'latitude': float(loc[0]), 'longitude': float(loc[1]),
                                                'altitude': float(loc[3])},
                usevegdem=1, anisotropic_sky=1, patch_option=2, cyl=1)
    prep = model.prepare_model(**args)
    z = np.zeros((rows, cols))
    state = dict(Tgmap1=z, Tgmap1E=z, Tgmap1S=z, Tgmap1W=z, Tgmap1N=z, TgOut1=z)
    CI, fdt, timeadd = 1., 1., 0.
    dectime = s['dectime']
    hours = list(golden[prefix + 'step_hours'])
    psi = np.where((met[:, 1] > 97) & (met[:, 1] < 300), 0.03, 0.5)
    worst = {}
    for i in range(met.shape[0]):
        if dectime[i] - np.floor(dectime[i]) == 0:
            CI = 1.
        r = model.Solweig_2022a_calc(
            i, p['dsm'], 1.0, rows, cols, p['svf'], p['svfN'], p['svfW'], p['svfE'], p['svfS'], p['svfveg'],
            p['svfNveg'], p['svfEveg'], p['svfSveg'], p['svfWveg'], p['svfaveg'], p['svfEaveg'], p['svfSaveg'],
            p['svfWaveg'], p['svfNaveg'], p['vegdem'], p['vegdem2'], 0.2, 0.7, 0.95, 0.9, 0.22, 0.06, 0.28,
            s['altitude'][i], s['azimuth'][i], s['zen'][i], s['jday'][i], 1, 1, p['buildings'], args['location'],
            psi[i], p['landcover'], p['lc_grid'], dectime[i], s['altmax'][i], p['dirwalls'], p['walls'], 1, 0,
            met[i, 11], met[i, 10], met[i, 14], met[i, 21], met[i, 22], met[i, 12], p['amaxvalue'], p['bush'],
            None, p['TgK'], p['Tstart'], p['alb_grid'], p['emis_grid'], p['TgK_wall'], p['Tstart_wall'],
            p['TmaxLST'], p['TmaxLST_wall'], 1.0, 22.0, p['svfalfa'], p['svfbuveg'], fdt, timeadd,
            dectime[1] - dectime[0], state['Tgmap1'], state['Tgmap1E'], state['Tgmap1S'], state['Tgmap1W'],
            state['Tgmap1N'], CI, state['TgOut1'], p['diffsh'], p['shmat'], p['vegshmat'], p['vbshvegshmat'],
            1, p['asvf'], 2, prepared=prep)
        out = dict(zip(model.OUTPUT_NAMES, r))
        CI, fdt, timeadd = out['CI'], out['firstdaytime'], out['timeadd']
        state = {k: out[k] for k in state}
        tm = out['Tmrt'].get()
        err = float(np.max(np.abs(tm - golden[prefix + 'step_Tmrt'][i])))
        worst['Tmrt'] = max(worst.get('Tmrt', 0), err)
        if i in hours:
            j = hours.index(i)
            for k in STEP_KEYS[1:]:
                got = out[k].get() if hasattr(out[k], 'get') else out[k]
                ref = golden[prefix + 'step_' + k][j].astype(np.float64)
                e = float(np.max(np.abs(got - ref) / (1.0 + np.abs(ref))))
                worst[k] = max(worst.get(k, 0), e)
    print('  worst:', {k: f'{v:.2e}' for k, v in worst.items()}, flush=True)
    assert worst['Tmrt'] < 0.05
    for k, v in worst.items():
        if k != 'Tmrt':
            assert v < 2e-3, (k, v)


def test_fused_wedge_against_the_fp64_oracle(golden):
    """The pixel pass evaluates the cylinder wedge F_sh in FP32 from y = tan(svfalfa) tan(zen)
    (umep_solweig.py:383-426); against the FP64 oracle over svfalfa from 0 to pi/2 and zeniths
    from 2 to 88 degrees the difference stays below 1e-6 (y = 0 gives the reference's NaN -> 0.5)."""
    from solweig_amd import model, runtime
    from oracle import umep
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    p = static_from_golden(golden, 'dayA_')
    rows, cols = p['dsm'].shape
    sweep = np.linspace(0.0, np.pi / 2, rows * cols).reshape(rows, cols)
    sweep.flat[1:40] = np.geomspace(1e-6, 1e-2, 39)
    args = dict(p)
    args.update(svfalfa=sweep, rows=rows, cols=cols, location={'latitude': 53.35}, usevegdem=1,
                anisotropic_sky=1, patch_option=2, cyl=1)
    prep = model.prepare_model(**args)
    k = runtime.FULL_PLANES.index('F_sh')
    worst = 0.0
    for zen_deg in (2.0, 15.0, 35.0, 60.0, 80.0, 88.0):
        print(f'fused wedge: zenith {zen_deg}', flush=True)
        model.run_step(prep, altitude=90.0 - zen_deg, azimuth=180.0, zen=np.radians(zen_deg), jday=172.0,
                       dectime=172.5, altmax=60.0, Ta=20.0, RH=60.0, radG=600.0, P=101.3, psi=0.03, CI=1.0,
                       firstdaytime=1, timeadd=0.0, timestepdec=1 / 24, lags=[0.0] * 6)
        prep.stream.sync()
        got = prep.full.plane(k, (rows, cols)).to_host().astype(np.float64)
        with np.errstate(all='ignore'):
            ref = umep.cylindric_wedge(np.radians(zen_deg), sweep)
        ref = np.where(np.isnan(ref), 0.5, ref)
        worst = max(worst, float(np.max(np.abs(got - ref))))
    print(f'  max |F_sh - FP64 oracle| = {worst:.2e}', flush=True)
    assert worst < 1e-6
