"""The per-step sub-models of model.py as standalone device calls (GPU), against the oracle:
cast_shadow bit-exact against the NumPy shadow march; the elementwise sub-models in FP64 to
round-off and in FP32 to float precision."""
import numpy as np
import pytest

from oracle import day as oday, umep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


@pytest.fixture(scope='module')
def prep():
    from solweig_amd import synthetic
    d = synthetic.small_domain(seed=17, size=38)
    return d, oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude)


@pytest.mark.parametrize('azimuth,altitude', [(95.3, 7.5), (180.0, 35.0), (251.7, 21.2), (333.0, 62.0),
                                              (45.0, 3.1)])
def test_cast_shadow_matches_oracle(rt, prep, azimuth, altitude):
    from solweig_amd import model
    d, p = prep
    print(f'cast_shadow azimuth {azimuth} altitude {altitude}', flush=True)
    aspect = p['dirwalls'].astype(np.float32).astype(np.float64) * np.pi / 180.0
    g = model.prepare_shadow(p['dsm'], p['walls'], aspect, 1.0, vegdem=p['vegdem'], vegdem2=p['vegdem2'],
                             bush=p['bush'], amaxvalue=p['amaxvalue'])
    dev = [v.to_host() for v in model.cast_shadow(g, azimuth, altitude)]
    ref = umep.shadows(p['dsm'], p['vegdem'], p['vegdem2'], azimuth, altitude, 1.0, p['amaxvalue'], p['bush'],
                       p['walls'], aspect)
    for k, (a, b) in enumerate(zip(dev, ref)):
        np.testing.assert_array_equal(a, b, err_msg=f'plane {k}')


def test_cast_shadow_buildings_only(rt, prep):
    from solweig_amd import model
    d, p = prep
    aspect = p['dirwalls'].astype(np.float32).astype(np.float64) * np.pi / 180.0
    g = model.prepare_shadow(p['dsm'], p['walls'], aspect, 1.0)
    dev = [v.to_host() for v in model.cast_shadow(g, 140.0, 18.0)]
    zero = np.zeros_like(p['dsm'])
    ref = umep.shadows(p['dsm'], zero, zero, 140.0, 18.0, 1.0, float(p['dsm'].max()), zero, p['walls'], aspect)
    for a, k in zip(dev, (1, 3, 4, 6, 7)):
        np.testing.assert_array_equal(a, ref[k], err_msg=f'plane {k}')


def _planes(rng, n, shape, lo, hi):
    return [rng.uniform(lo, hi, size=shape) for _ in range(n)]


@pytest.mark.parametrize('dtype,rtol', [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_elementwise_submodels_match_oracle(rt, dtype, rtol):
    from solweig_amd import flux
    rng = np.random.default_rng(9)
    shape = (13, 17)
    cast = lambda x: np.asarray(x, dtype=dtype).astype(np.float64)

    svfalfa = rng.uniform(0.05, 1.4, size=shape)
    w = flux.cylindric_wedge(0.7, svfalfa, *shape, dtype=dtype).to_host()
    ref = umep.cylindric_wedge(0.7, svfalfa)
    assert np.array_equal(np.isnan(w), np.isnan(ref))
    np.testing.assert_allclose(w[~np.isnan(ref)], ref[~np.isnan(ref)], rtol=rtol)

    cur, state = _planes(rng, 2, shape, 300, 600)
    for first, timeadd in ((1, 0.0), (0, 1 / 24), (0, 10 / 1440)):
        d, nxt, s = flux.TsWaveDelay_2015a(cast(cur), first, timeadd, 1 / 24, cast(state), dtype=dtype)
        rd, rnxt, rs = umep.wave_delay(cast(cur), first, timeadd, 1 / 24, cast(state))
        np.testing.assert_allclose(d.to_host(), rd, rtol=rtol)
        np.testing.assert_allclose(s.to_host(), rs, rtol=rtol)
        assert nxt == rnxt

    g = [cast(x) for x in _planes(rng, 12, shape, 0.05, 0.9)]
    out = flux.Kup_veg_2015a(610.0, 140.0, 700.0, 38.0, g[0], 0.2, g[1], *g[2:], dtype=dtype)
    ref = umep.kup(610.0, 140.0, 700.0, 38.0, g[0], 0.2, g[1], *g[2:])
    for a, b in zip(out, ref):
        np.testing.assert_allclose(a.to_host(), b, rtol=rtol)

    k = [cast(x) for x in _planes(rng, 15, shape, 20, 500)]
    for cyl, ani in ((1, 1), (1, 0), (0, 0)):
        t, s = flux.assemble_tmrt(*k, 0.7, 0.95, 0.22, 0.06, 0.28, cyl, ani, dtype=dtype)
        rt_, rs = umep.assemble_tmrt(*k, 0.7, 0.95, 0.22, 0.06, 0.28, cyl, ani)
        np.testing.assert_allclose(s.to_host(), rs, rtol=rtol)
        np.testing.assert_allclose(t.to_host(), rt_, rtol=rtol, atol=1e-9 if dtype == np.float64 else 2e-3)

    TgK, Tstart = cast(rng.uniform(0.2, 0.6, shape)), cast(rng.uniform(-9, -3, shape))
    TmaxLST = cast(rng.uniform(13, 16, shape))
    for dec, lc in ((172 + 13 / 24, 0), (172 + 3 / 24, 1), (172 + 9 / 24, 1)):
        tg, tw = flux.surface_temperature(TgK, Tstart, 0.37, -3.41, TmaxLST, 15.0, 58.0, dec, 4.9, 0.93, lc, dtype=dtype)
        rtg, rtw = umep.surface_temperature(TgK, Tstart, 0.37, -3.41, TmaxLST, 15.0, 58.0, dec, 4.9, 0.93, lc)
        np.testing.assert_allclose(tg.to_host(), rtg, rtol=rtol, atol=1e-10 if dtype == np.float64 else 1e-5)
        np.testing.assert_allclose(tw.to_host(), np.broadcast_to(rtw, shape), rtol=rtol, atol=1e-10)


@pytest.mark.parametrize('case', range(4))
def test_lside_matches_umep_golden(rt, case):
    from solweig_amd import flux
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'umep_lside_golden.npz'))
    from test_oracle_submodels import lside_case
    args, ref = lside_case(g, case)
    for ani in (0, 1):
        out = np.stack([v.to_host() for v in flux.Lside_veg_v2022a(*args, ani)])
        expect = ref if ani == 0 else np.stack(umep.lside(*args, 1))
        np.testing.assert_allclose(out, expect, rtol=1e-12, atol=1e-9)


def _gvf_fields(seed, rows, cols, landcover):
    rng = np.random.default_rng(seed)
    u = lambda: rng.random((rows, cols))
    buildings = (u() < 0.8).astype(np.float64)
    buildings[u() < 0.1] = 0.5                     # a ground mask of any value: the walk's running minimum
    walls = np.where(u() < 0.3, rng.uniform(1.0, 12.0, (rows, cols)), 0.0)
    wallsun = np.where(u() < 0.6, walls, walls * 0.3)
    dirwalls = rng.uniform(0.0, 360.0, (rows, cols))
    shadow = u()
    shadow[u() < 0.3] = 1.0
    shadow[u() < 0.2] = 0.0
    f = dict(buildings=buildings, walls=walls, wallsun=wallsun, dirwalls=dirwalls, shadow=shadow,
             Tg=rng.uniform(-2.0, 15.0, (rows, cols)), emis=rng.uniform(0.9, 0.98, (rows, cols)),
             alb=rng.uniform(0.1, 0.3, (rows, cols)), Ta=23.4)
    if landcover:
        f.update(Tgwall=rng.uniform(0.0, 9.0, (rows, cols)), lc=rng.choice([1.0, 2.0, 3.0, 5.0, 6.0, 7.0], (rows, cols)),
                 Twater=17.9)
    else:
        f.update(Tgwall=4.2, lc=np.zeros((rows, cols)), Twater=None)
    return f


@pytest.mark.parametrize('landcover', [0, 1])
@pytest.mark.parametrize('scale,first,second', [(1.0, 1.0, 22.0), (2.0, 1.0, 6.0)])
def test_gvf_matches_oracle(rt, scale, first, second, landcover):
    """Standalone ground view factors (17 planes + the water-replaced Tg) on a 23 x 31 raster of
    random fields, a fractional shadow and a ground mask with 0.5 values; land cover with water
    pixels; a scale of 2 makes the near field two steps long."""
    from solweig_amd import model
    print(f'gvf scale {scale} first {first} second {second} landcover {landcover}', flush=True)
    f = _gvf_fields(5 + landcover, 23, 31, landcover)
    geo = model.prepare_gvf_geometry(f['walls'], f['buildings'], f['dirwalls'], scale, first, second)
    tg_ref = f['Tg'].copy()
    ref = umep.gvf(f['wallsun'], f['walls'], f['buildings'], scale, f['shadow'], first, second, f['dirwalls'],
                   tg_ref, f['Tgwall'], f['Ta'], f['emis'], 0.9, f['alb'], 0.2, f['Twater'], f['lc'], landcover)
    ref = np.stack(ref)
    for dtype, rtol, atol in ((np.float64, 1e-11, 1e-9), (np.float32, 5e-5, 2e-4)):
        res = model.gvf(geo, f['wallsun'], f['shadow'], f['Tg'], f['Tgwall'], f['Ta'], f['emis'], 0.9, f['alb'],
                        umep.SBC, 0.2, f['Twater'], f['lc'], landcover, dtype=dtype)
        got = np.stack([v.to_host() for v in res.outputs]).astype(np.float64)
        for k in range(17):
            np.testing.assert_allclose(got[k], ref[k], rtol=rtol, atol=atol, err_msg=f'{np.dtype(dtype)} plane {k}')
        np.testing.assert_allclose(res.surface_temperature.to_host(), tg_ref, rtol=rtol, atol=1e-6)


@pytest.mark.parametrize('prefix,hour', [('dayA_', 8), ('dayA_', 14), ('dayB_', 10), ('dayC_', 12)])
def test_gvf_matches_umep_golden(rt, golden, prefix, hour):
    """The ground view factors UMEP wrote for three domains (walls, land cover) against the device."""
    from solweig_amd import model
    from golden_inputs import static_from_golden
    p = static_from_golden(golden, prefix)
    met = golden[prefix + 'met']
    sh = golden[prefix + f'shadow{hour}'].astype(np.float64)
    shadow = golden[prefix + f'gvf{hour}_shadow'].astype(np.float64)
    tg = golden[prefix + f'gvf{hour}_Tg'].astype(np.float64)
    twater = float(np.mean(met[:, 11])) if p['landcover'] else None
    geo = model.prepare_gvf_geometry(p['walls'], p['buildings'], p['dirwalls'], 1.0, 1.0, 22.0)
    res = model.gvf(geo, sh[4], shadow, tg, 3.5, float(met[hour, 11]), p['emis_grid'], 0.9, p['alb_grid'],
                    umep.SBC, 0.2, twater, p['lc_grid'], p['landcover'])
    got = np.stack([v.to_host() for v in res.outputs])
    np.testing.assert_allclose(got, golden[prefix + f'gvf{hour}'].astype(np.float64), rtol=2e-6, atol=1e-4)
    ref = np.stack(umep.gvf(sh[4], p['walls'], p['buildings'], 1.0, shadow, 1.0, 22.0, p['dirwalls'], tg.copy(),
                            3.5, float(met[hour, 11]), p['emis_grid'], 0.9, p['alb_grid'], 0.2, twater,
                            p['lc_grid'], p['landcover']))
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-9)


def _aniso():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), 'golden', 'umep_aniso_golden.npz'))


@pytest.mark.parametrize('case', range(3))
def test_longwave_matches_umep_golden(rt, case):
    """Anisotropic longwave over random masks (day patches from Perez, a night step) against the
    vectors UMEP's Lcyl_v2022a wrote and the FP64 oracle; masks as uint8 and as float32."""
    from solweig_amd import model
    from test_oracle_submodels import lcyl_case
    g = _aniso()
    args, ref = lcyl_case(g, case)
    esky, patches, Ta, Tgwall, ewall, Lup, sh, veg, vb, salt, sazi, asvf = args
    masks = g[f'l{case}_masks']
    for mk in (masks, masks.astype(np.float32)):
        geo = model.prepare_geometry(patches, *mk)
        for dtype, rtol, atol in ((np.float64, 1e-12, 1e-9), (np.float32, 2e-5, 2e-3)):
            out = np.stack([v.to_host() for v in model.longwave(geo, esky, Ta, Tgwall, ewall, Lup, salt, sazi, asvf,
                                                                dtype=dtype)])
            np.testing.assert_allclose(out, ref, rtol=rtol, atol=atol, err_msg=str(np.dtype(dtype)))
    tgw = np.random.default_rng(case).uniform(0.0, 9.0, Lup.shape)
    oref = np.stack(umep.lcyl(esky, patches, Ta, tgw, ewall, Lup, sh, veg, vb, salt, sazi, asvf))
    out = np.stack([v.to_host() for v in model.longwave(geo, esky, Ta, tgw, ewall, Lup, salt, sazi, asvf)])
    np.testing.assert_allclose(out, oref, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize('case', range(5))
def test_kside_matches_umep_golden(rt, case):
    """Kside_veg_v2022a in the cylinder / box x anisotropic / weighted branches against UMEP's
    vectors (FP64 and FP32), with and without a prepared shortwave state."""
    from solweig_amd import model
    from test_oracle_submodels import kside_case
    g = _aniso()
    args, ref = kside_case(g, case)
    (radI, radD, radG, shadow, svfS, svfW, svfN, svfE, svfEveg, svfSveg, svfWveg, svfNveg, az, alt, psi, t, albedo,
     F_sh, KupE, KupS, KupW, KupN, cyl, lv, ani, diffsh, asvf, sh, veg, vb) = args
    rows, cols = shadow.shape
    print(f'kside case {case}: cyl {cyl} anisotropic {ani}', flush=True)
    for dtype, rtol, atol in ((np.float64, 1e-12, 1e-9), (np.float32, 2e-5, 2e-3)):
        common = (radI, radD, radG, shadow, svfS, svfW, svfN, svfE, svfEveg, svfSveg, svfWveg, svfNveg, az, alt, psi,
                  t, albedo, F_sh, KupE, KupS, KupW, KupN, cyl, lv, ani, diffsh, rows, cols, asvf, sh, veg, vb)
        out = np.stack([v.to_host() for v in model.Kside_veg_v2022a(*common, dtype=dtype)])
        np.testing.assert_allclose(out, ref, rtol=rtol, atol=atol, err_msg=str(np.dtype(dtype)))
        if ani == 1:
            prep = model.prepare_shortwave(model.prepare_geometry(lv, sh, veg, vb), diffsh, dtype=dtype)
            again = np.stack([v.to_host() for v in model.Kside_veg_v2022a(*common, dtype=dtype, prepared=prep)])
            np.testing.assert_array_equal(again, out)
