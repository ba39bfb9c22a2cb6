"""The per-step drop-in (model.Solweig_2022a_calc) against UMEP's own step (golden vectors, GPU)."""
import numpy as np
import pytest

from golden_inputs import static_from_golden, sun_from_golden

pytestmark = pytest.mark.gpu

STEP_KEYS = ('Tmrt', 'Kdown', 'Kup', 'Ldown', 'Lup', 'Tg', 'shadow', 'Keast', 'Ksouth', 'Kwest',
             'Knorth', 'Least', 'Lsouth', 'Lwest', 'Lnorth', 'KsideI', 'TgOut', 'Lside', 'KsideD',
             'dRad', 'Kside')


def _golden_day(golden, prefix, beyond_storage=False, **prep_kw):
    """Run the drop-in step over a UMEP golden day; worst errors: Tmrt absolute (K), the other
    per-pixel outputs relative to 1 + |value| at the golden hours. The fixture holds UMEP's FP64
    outputs rounded to float32; with ``beyond_storage`` each error is first reduced by half a
    float32 ulp of the stored value (the part the storage rounding alone explains)."""
    from solweig_amd import model, runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    p = static_from_golden(golden, prefix)
    met = golden[prefix + 'met']
    s = sun_from_golden(golden, prefix)
    loc = golden[prefix + 'loc']
    rows, cols = p['dsm'].shape
    args = dict(p)
    args.update(rows=rows, cols=cols, location={'latitude': float(loc[0]), 'longitude': float(loc[1]),
                                                'altitude': float(loc[3])},
                usevegdem=1, anisotropic_sky=1, patch_option=2, cyl=1)
    prep = model.prepare_model(**args, **prep_kw)
    z = np.zeros((rows, cols))
    state = dict(Tgmap1=z, Tgmap1E=z, Tgmap1S=z, Tgmap1W=z, Tgmap1N=z, TgOut1=z)
    CI, fdt, timeadd = 1., 1., 0.
    dectime = s['dectime']
    hours = list(golden[prefix + 'step_hours'])
    psi = np.where((met[:, 1] > 97) & (met[:, 1] < 300), 0.03, 0.5)
    twater = float(np.mean(met[:, 11])) if p['landcover'] else None
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
            twater, p['TgK'], p['Tstart'], p['alb_grid'], p['emis_grid'], p['TgK_wall'], p['Tstart_wall'],
            p['TmaxLST'], p['TmaxLST_wall'], 1.0, 22.0, p['svfalfa'], p['svfbuveg'], fdt, timeadd,
            dectime[1] - dectime[0], state['Tgmap1'], state['Tgmap1E'], state['Tgmap1S'], state['Tgmap1W'],
            state['Tgmap1N'], CI, state['TgOut1'], p['diffsh'], p['shmat'], p['vegshmat'], p['vbshvegshmat'],
            1, p['asvf'], 2, prepared=prep)
        out = dict(zip(model.OUTPUT_NAMES, r))
        CI, fdt, timeadd = out['CI'], out['firstdaytime'], out['timeadd']
        state = {k: out[k] for k in state}
        tm = out['Tmrt'].get()
        ref = golden[prefix + 'step_Tmrt'][i]
        err = float(np.max(_excess(tm, ref, beyond_storage)))
        worst['Tmrt'] = max(worst.get('Tmrt', 0), err)
        if i in hours:
            j = hours.index(i)
            for k in STEP_KEYS[1:]:
                got = out[k].get() if hasattr(out[k], 'get') else out[k]
                ref = golden[prefix + 'step_' + k][j]
                e = float(np.max(_excess(got, ref, beyond_storage) / (1.0 + np.abs(ref.astype(np.float64)))))
                worst[k] = max(worst.get(k, 0), e)
    return worst


def _excess(got, ref32, beyond_storage):
    d = np.abs(np.asarray(got, np.float64) - ref32.astype(np.float64))
    if beyond_storage:
        d = np.maximum(d - 0.5 * np.spacing(np.abs(ref32)).astype(np.float64), 0.0)
    return d


@pytest.mark.parametrize('prefix', ['dayA_', 'dayB_', 'dayC_'])
def test_step_api_matches_umep(golden, prefix):
    print(f'step API (fused FP32) vs UMEP {prefix}', flush=True)
    worst = _golden_day(golden, prefix)
    print('  worst:', {k: f'{v:.2e}' for k, v in worst.items()}, flush=True)
    assert worst['Tmrt'] < 0.05
    for k, v in worst.items():
        if k != 'Tmrt':
            assert v < 2e-3, (k, v)


@pytest.mark.parametrize('prefix', ['dayA_', 'dayB_', 'dayC_'])
def test_stepwise_fp64_step_matches_umep(golden, prefix):
    """dtype=np.float64: the reference's sequence of sub-model calls in FP64 reproduces UMEP's own
    FP64 step to the fixture's float32 storage: beyond half a float32 ulp of each stored value,
    Tmrt stays within 1e-9 K and every other plane within 1e-10 relative."""
    import numpy
    print(f'step API (stepwise FP64) vs UMEP {prefix}', flush=True)
    worst = _golden_day(golden, prefix, beyond_storage=True, dtype=numpy.float64)
    print('  worst beyond storage rounding:', {k: f'{v:.2e}' for k, v in worst.items()}, flush=True)
    assert worst['Tmrt'] < 1e-9
    for k, v in worst.items():
        if k != 'Tmrt':
            assert v < 1e-10, (k, v)


@pytest.mark.parametrize('prefix', ['dayA_', 'dayC_'])
def test_stepwise_fp32_step_matches_umep(golden, prefix):
    """prepare_model(stepwise=True): the same sequence in FP32 (the plain per-step forward of the
    reference's structure) stays within the FP32 bar."""
    print(f'step API (stepwise FP32) vs UMEP {prefix}', flush=True)
    worst = _golden_day(golden, prefix, stepwise=True)
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


PLANES = ('Tmrt', 'Kdown', 'Kup', 'Ldown', 'Lup', 'Tg', 'shadow', 'Keast', 'Ksouth', 'Kwest', 'Knorth', 'Least',
          'Lsouth', 'Lwest', 'Lnorth', 'KsideI', 'TgOut', 'Lside', 'KsideD', 'dRad', 'Kside')
# FP32 bars: Tmrt (and UTCI) in kelvin; W m-2 planes relative to 1 + |value|
BAR_K, BAR_REL = 0.05, 2e-3


def test_fp32_error_per_output_against_the_fp64_step(golden):
    """The error statement per output: one day of a 128 x 128 C1-generator domain through the fused
    FP32 passes (the batched engine's Tmrt and UTCI, and every per-pixel plane of the full-mode
    drop-in step) against the stepwise FP64 drop-in, itself pinned to UMEP's FP64 step."""
    from solweig_amd import batched, model, runtime, static, synthetic, sun as psun
    from oracle import utci as outci
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    d = synthetic.make_c1(seed=11, size=128)
    s = psun.met_sun(d.met, d.location, d.utc)
    pd = static.PreparedDomain(d, keep_host=True)
    inputs = static.model_inputs(pd)
    f32 = model.prepare_model(**inputs)
    f64 = model.prepare_model(dtype=np.float64, **inputs)
    eng = batched.Engine([pd], n_hours=24)
    eng.load(eng.tables(sun=[s]))
    eng.reset_state()
    eng.forward_day()
    tm_b, ut_b = eng.outputs(0)
    ground = inputs['buildings'] == 1
    z = np.zeros(pd.shape)
    st = {p: dict(Tgmap1=z, Tgmap1E=z, Tgmap1S=z, Tgmap1W=z, Tgmap1N=z, TgOut1=z, CI=1., fdt=1., timeadd=0.)
          for p in (32, 64)}
    err = {k: [0.0, 0.0, 0] for k in PLANES + ('UTCI (batched)', 'Tmrt (batched)')}
    dec = s['dectime']
    for i in range(24):
        out = {}
        for bits, prep in ((32, f32), (64, f64)):
            q = st[bits]
            if dec[i] - np.floor(dec[i]) == 0:
                q['CI'] = 1.
            r = dict(zip(model.OUTPUT_NAMES, model.Solweig_2022a_calc(
                i, inputs['dsm'], 1.0, *pd.shape, *(inputs[k] for k in (
                    'svf', 'svfN', 'svfW', 'svfE', 'svfS', 'svfveg', 'svfNveg', 'svfEveg', 'svfSveg', 'svfWveg',
                    'svfaveg', 'svfEaveg', 'svfSaveg', 'svfWaveg', 'svfNaveg', 'vegdem', 'vegdem2')),
                0.2, 0.7, 0.95, 0.9, 0.22, 0.06, 0.28, s['altitude'][i], s['azimuth'][i], s['zen'][i], s['jday'][i],
                1, 1, inputs['buildings'], inputs['location'], 0.03, 0, inputs['lc_grid'], dec[i], s['altmax'][i],
                inputs['dirwalls'], inputs['walls'], 1, 0, d.met[i, 11], d.met[i, 10], d.met[i, 14], d.met[i, 21],
                d.met[i, 22], d.met[i, 12], inputs['amaxvalue'], inputs['bush'], None, inputs['TgK'],
                inputs['Tstart'], inputs['alb_grid'], inputs['emis_grid'], inputs['TgK_wall'],
                inputs['Tstart_wall'], inputs['TmaxLST'], inputs['TmaxLST_wall'], 1.0, 22.0, inputs['svfalfa'],
                inputs['svfbuveg'], q['fdt'], q['timeadd'], 1 / 24, q['Tgmap1'], q['Tgmap1E'], q['Tgmap1S'],
                q['Tgmap1W'], q['Tgmap1N'], q['CI'], q['TgOut1'], inputs['diffsh'], inputs['shmat'],
                inputs['vegshmat'], inputs['vbshvegshmat'], 1, inputs['asvf'], 2, prepared=prep)))
            q.update(CI=r['CI'], fdt=r['firstdaytime'], timeadd=r['timeadd'],
                     **{k: r[k] for k in ('Tgmap1', 'Tgmap1E', 'Tgmap1S', 'Tgmap1W', 'Tgmap1N', 'TgOut1')})
            out[bits] = {k: (r[k].get() if hasattr(r[k], 'get') else np.asarray(r[k])).astype(np.float64)
                         for k in PLANES}
        ref_tm = out[64]['Tmrt']
        ref_ut = np.where(ground, outci.utci(np.float32(d.met[i, 11]), np.float32(d.met[i, 10]), ref_tm,
                                             np.float32(d.met[i, 9])), np.nan)
        pairs = [(k, out[32][k], out[64][k]) for k in PLANES] + \
                [('Tmrt (batched)', tm_b[i], ref_tm), ('UTCI (batched)', ut_b[i][ground], ref_ut[ground])]
        for k, a, b in pairs:
            e = np.abs(np.asarray(a, np.float64) - b)
            if k not in ('Tmrt', 'Tmrt (batched)', 'UTCI (batched)'):
                e = e / (1.0 + np.abs(b))
            err[k][0] = max(err[k][0], float(e.max()))
            err[k][1] += float((e ** 2).sum())
            err[k][2] += e.size
    print('  per-output error of the FP32 passes vs the FP64 step (max, RMS; K for Tmrt/UTCI, '
          'relative to 1 + |value| otherwise):', flush=True)
    for k, (mx, sq, n) in err.items():
        print(f'    {k:16s} {mx:.2e} {np.sqrt(sq / n):.2e}', flush=True)
    for k, (mx, sq, n) in err.items():
        assert mx < (BAR_K if k.startswith(('Tmrt', 'UTCI')) else BAR_REL), (k, mx)
