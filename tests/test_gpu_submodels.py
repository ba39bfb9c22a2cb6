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
