"""The oracle against vectors written by the reference's UMEP module (tests/golden)."""
import numpy as np
import pytest

from oracle import day as oday, spa, svf as osvf, umep, utci as outci
from golden_inputs import static_from_golden, sun_from_golden, unpack

from solweig_amd import sun as psun, synthetic


@pytest.mark.parametrize('k', range(5))
def test_metdata_matches_umep(golden, k):
    met = np.full((24, 24), -999.0)
    met[:, :4] = golden[f'met{k}_in']
    lat, lon, utc, alt = golden[f'met{k}_loc']
    loc = {'latitude': lat, 'longitude': lon, 'altitude': alt}
    ref = golden[f'met{k}_out']
    o = spa.metdata(met, loc, utc)
    p = psun.met_sun(met, loc, utc)
    for i, key in enumerate(('altitude', 'azimuth', 'zen', 'jday', 'dectime', 'altmax')):
        np.testing.assert_allclose(o[key], ref[i], rtol=0, atol=1e-9, err_msg=key)
        np.testing.assert_allclose(p[key], ref[i], rtol=0, atol=1e-9, err_msg=key)


def test_clearness_matches_umep(golden):
    for row in golden['clearness']:
        zen, jday, Ta, RH, radG, lat, P = row[:7]
        got = umep.clearness_index(zen, jday, Ta, RH, radG, lat, P)
        np.testing.assert_allclose(np.array(got, dtype=float), row[7:], rtol=1e-13)


def test_perez_matches_umep(golden):
    for row in golden['perez']:
        zen, azi, radD, radI, jday, clear, bright = row[:7]
        lv, c, b = umep.perez(zen, azi, radD, radI, jday, 2)
        np.testing.assert_allclose(lv[:, 2], row[7:], rtol=1e-12)
        assert c == pytest.approx(clear, rel=1e-14)
    np.testing.assert_array_equal(lv[:, :2], golden['perez_angles'])


def test_perez_luminance_normalised():
    lv, _, _ = umep.perez(40.0, 180.0, 120.0, 500.0, 172, 2)
    assert lv[:, 2].sum() == pytest.approx(1.0, abs=1e-14)


def test_utci_matches_reference_expression(golden):
    Ta, RH, Tmrt, va, ref = golden['utci']
    got = outci.utci(Ta, RH, Tmrt, va)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9)


def test_utci_published_example():
    # Broede et al. (2012) reference implementation example, as listed by pythermalcomfort:
    # Ta = Tmrt = 25 C, va = 1 m/s, RH = 50 % -> UTCI 24.6 C (one decimal).
    assert round(float(outci.utci(25.0, 50.0, 25.0, 1.0)), 1) == 24.6


def test_utci_factored_form_equals_full():
    rng = np.random.default_rng(3)
    for _ in range(20):
        Ta, RH, va = rng.uniform(-10, 40), rng.uniform(10, 100), rng.uniform(0.5, 10)
        c = outci.coefficients(Ta, RH, va)
        d = rng.uniform(-20, 50, 50)
        full = outci.utci(Ta, RH, Ta + d, va)
        fact = sum(c[k] * d ** k for k in range(7))
        np.testing.assert_allclose(fact, full, rtol=0, atol=1e-9)


def _gvf_case(golden, prefix, hour):
    p = static_from_golden(golden, prefix)
    met = golden[prefix + 'met']
    sh = golden[prefix + f'shadow{hour}'].astype(np.float64)
    shadow = golden[prefix + f'gvf{hour}_shadow'].astype(np.float64)
    tg = golden[prefix + f'gvf{hour}_Tg'].astype(np.float64)
    got = umep.gvf(sh[4], p['walls'], p['buildings'], 1.0, shadow, 1.0, 22.0, p['dirwalls'], tg.copy(),
                   3.5, float(met[hour, 11]), p['emis_grid'], 0.9, p['alb_grid'], 0.2,
                   float(np.mean(met[:, 11])) if p['landcover'] else None, p['lc_grid'], p['landcover'])
    return np.stack(got), golden[prefix + f'gvf{hour}'].astype(np.float64)


@pytest.mark.parametrize('prefix,hour', [('dayA_', 8), ('dayA_', 14), ('dayB_', 10), ('dayC_', 12)])
def test_gvf_matches_umep(golden, prefix, hour):
    got, ref = _gvf_case(golden, prefix, hour)
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-4)


@pytest.mark.parametrize('prefix,hour', [('dayA_', 8), ('dayA_', 14), ('dayB_', 10), ('dayC_', 12)])
def test_shadows_match_umep(golden, prefix, hour):
    p = static_from_golden(golden, prefix)
    s = sun_from_golden(golden, prefix)
    got = umep.shadows(p['dsm'], p['vegdem'], p['vegdem2'], s['azimuth'][hour], s['altitude'][hour], 1.0,
                       p['amaxvalue'], p['bush'], p['walls'], p['dirwalls'] * np.pi / 180.)
    ref = golden[prefix + f'shadow{hour}'].astype(np.float64)
    for k in range(3):
        np.testing.assert_array_equal(got[k], ref[k])
    np.testing.assert_allclose(np.stack(got[3:]), ref[3:], rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize('prefix', ['dayA_', 'dayB_', 'dayC_'])
def test_day_matches_umep(golden, prefix):
    p = static_from_golden(golden, prefix)
    met = golden[prefix + 'met']
    s = sun_from_golden(golden, prefix)
    keys = ('Tmrt', 'Kdown', 'Kup', 'Ldown', 'Lup', 'Lside', 'Kside', 'KsideD', 'dRad', 'shadow',
            'Least', 'Lnorth', 'Keast', 'TgOut')
    out = oday.run_day(p, met, s, keep=keys)
    hours = golden[prefix + 'step_hours']
    tm = np.stack([o['Tmrt'] for o in out])
    np.testing.assert_allclose(tm, golden[prefix + 'step_Tmrt'], rtol=0, atol=2e-5)
    for k in keys[1:]:
        got = np.stack([out[h][k] for h in hours])
        ref = golden[prefix + 'step_' + k].astype(np.float64)
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-4, err_msg=k)


def test_brute_force_pins_shadow_march():
    d = synthetic.small_domain(seed=31, size=20, boxes=3)
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude,
                     svf=tuple(np.ones((20, 20)) for _ in range(15)) + (np.ones((20, 20, 153), np.float32),) * 3 + (None,))
    for az, al in ((135.0, 30.0), (250.0, 12.0), (10.0, 55.0), (300.0, 8.0), (90.0, 20.0)):
        v = umep.shadows(p['dsm'], p['vegdem'], p['vegdem2'], az, al, 1.0, p['amaxvalue'], p['bush'],
                         p['walls'], p['dirwalls'] * np.pi / 180.)
        from oracle import brute
        b = brute.shadow_masks(p['dsm'], p['vegdem'], p['vegdem2'], p['bush'], az, al, 1.0, p['amaxvalue'])
        for k in range(3):
            np.testing.assert_array_equal(v[k], b[k])


def test_box_shadow_closed_form():
    # One box of height h on flat ground, sun due south (azimuth 180): the ray steps straight
    # along +row, so the ground cell d rows north of the box is shaded iff d*tan(alt) < h.
    a = np.zeros((60, 20))
    a[40:44, 8:12] = 10.0
    z = np.zeros_like(a)
    alt = 35.0
    v = umep.shadows(a, z, z, 180.0, alt, 1.0, 10.0, z, z, z)
    sh = v[1]
    for d in range(1, 40):
        expect = 0.0 if d * np.tan(np.radians(alt)) < 10.0 else 1.0
        assert sh[40 - d, 10] == expect, d
    assert (sh[45:, :] == 1).all()


def test_flat_ground_svf_is_one():
    a = np.zeros((12, 12))
    out = osvf.svf_calculator(0.0, a, a, a, a, 1.0)
    # svf, svfaveg and svfveg integrate the annulus weights over the whole vault: exactly 1.
    for plane in (out[0], out[1], out[11]):
        np.testing.assert_allclose(plane, 1.0, rtol=0, atol=1e-12)
    w = osvf.ray_tables(4, 4, 1.0, 1.0)['weights']
    assert w[:, 0] @ np.array(osvf.COUNTS) == pytest.approx(1.0, abs=1e-12)


def test_wave_delay_closed_form():
    x = np.array([3.0])
    state = np.array([1.0])
    out, t, st = umep.wave_delay(x, 0, 1 / 24, 1 / 24, state)
    w = np.exp(-33.27 / 24)
    assert out[0] == pytest.approx(3.0 * (1 - w) + 1.0 * w, rel=1e-15)
    assert st[0] == out[0] and t == 1 / 24


def test_open_sky_ldown_closed_form():
    # Fully visible sky: Ldown is the Martin-Berdahl band emissivities weighted over the vault.
    n = (4, 4, 153)
    ones = np.ones(n)
    lv = np.column_stack(umep.create_patches(2) + (np.zeros(153),))
    Lup = np.full((4, 4), 400.0)
    Ld, Ls, *_ = umep.lcyl(0.8, lv, 15.0, 0.0, 0.9, Lup, ones, ones, ones, -5.0, 0.0, np.zeros((4, 4)))
    ster = umep.steradians(lv[:, 0])
    e = 1 - 0.2 * np.exp(0.308 * (1.7 - 1 / np.cos(np.radians(90 - lv[:, 0]))))
    expect = np.sum(e * umep.SBC * 288.15 ** 4 / np.pi * ster * np.sin(np.radians(lv[:, 0])))
    np.testing.assert_allclose(Ld, expect, rtol=1e-12)


def test_gvf_uniform_field_identity():
    # Flat open ground, uniform fields: every gathered sample equals the centre value, so the
    # GVF Lup is the Stefan-Boltzmann flux of the shaded/sunlit ground exactly.
    n = 50
    z = np.zeros((n, n))
    one = np.ones((n, n))
    shadow = np.full((n, n), 0.7)
    Tg = np.full((n, n), 6.0)
    out = umep.gvf(z, z, one, 1.0, shadow, 1.0, 22.0, z, Tg, 0.0, 20.0, one * 0.95, 0.9, one * 0.2, 0.2,
                   None, z, 0)
    c = slice(22, n - 22)
    flux = umep.SBC * 0.95 * (6.0 * 0.7 + 20.0 + 273.15) ** 4
    np.testing.assert_allclose(out[0][c, c], flux, rtol=1e-12)
    np.testing.assert_allclose(out[1][c, c], 0.2 * 0.7, rtol=1e-12)
    np.testing.assert_allclose(out[2][c, c], 0.2, rtol=1e-12)
    np.testing.assert_allclose(out[16][c, c], 0.7, rtol=1e-12)
