"""Oracle per-step sub-models (the drop-ins of model.py's sub-APIs) pinned on the CPU."""
import os

import numpy as np
import pytest

from oracle import umep

LSIDE = os.path.join(os.path.dirname(__file__), 'golden', 'umep_lside_golden.npz')
NAMES = ('svfS', 'svfW', 'svfN', 'svfE', 'svfEveg', 'svfSveg', 'svfWveg', 'svfNveg', 'svfEaveg', 'svfSaveg',
         'svfWaveg', 'svfNaveg')


def lside_case(g, c):
    p = f'c{c}_'
    az, alt, Ta, SBC, ewall, esky, t, CI = g[p + 'scalars']
    args = [g[p + n] for n in NAMES] + [az, alt, Ta, g[p + 'Tw'], SBC, ewall, g[p + 'Ldown'], esky, t, g[p + 'F_sh'],
                                        CI, g[p + 'LupE'], g[p + 'LupS'], g[p + 'LupW'], g[p + 'LupN']]
    return args, g[p + 'out']


@pytest.mark.parametrize('case', range(4))
def test_lside_isotropic_matches_umep(case):
    g = np.load(LSIDE)
    args, ref = lside_case(g, case)
    out = np.stack(umep.lside(*args, 0))
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-10)


def test_lside_anisotropic_is_half_lup():
    g = np.load(LSIDE)
    args, _ = lside_case(g, 0)
    out = umep.lside(*args, 1)
    for o, up in zip(out, args[-4:]):
        np.testing.assert_array_equal(o, up * 0.5)


def test_lside_weights_sum():
    rng = np.random.default_rng(3)
    svf, veg, aveg = rng.uniform(0.2, 0.9, (3, 50))
    wveg, wwall, wsky, wrefl = umep.lside_weights(svf, veg, aveg)
    np.testing.assert_allclose(wveg + wwall, wrefl, rtol=1e-14)    # veg + wall = everything not sky
    np.testing.assert_allclose(wsky + wrefl, 1.0, rtol=1e-14)


def test_surface_temperature_closed_form():
    TgK, Tstart, TmaxLST = np.array([0.37, 0.58]), np.array([-3.41, -9.78]), np.array([15.0, 15.0])
    SNUP, altmax, CI = 5.0, 58.0, 0.9
    at_rise = umep.surface_temperature(TgK, Tstart, 0.37, -3.41, TmaxLST, 15.0, altmax, 100 + SNUP / 24, SNUP, CI)
    np.testing.assert_allclose(at_rise[0], 0.0, atol=1e-12)
    at_max = umep.surface_temperature(TgK, Tstart, 0.37, -3.41, TmaxLST, 15.0, altmax, 100 + 15.0 / 24, SNUP, CI)
    np.testing.assert_allclose(at_max[0], (TgK * altmax + Tstart) * CI, rtol=1e-14)
    night = umep.surface_temperature(TgK, Tstart, 0.37, -3.41, TmaxLST, 15.0, altmax, 100 + 2 / 24, SNUP, CI, 1)
    assert np.all(night[0] == 0) and np.all(night[1] == 0)          # land cover clamps Tg, wall clamps always


def test_assemble_round_trip():
    rng = np.random.default_rng(5)
    f = rng.uniform(50, 450, (15, 20))
    Tmrt, S = umep.assemble_tmrt(*f, 0.7, 0.95, 0.22, 0.06, 0.28, 1, 1)
    np.testing.assert_allclose((Tmrt + 273.2) ** 4 * 0.95 * 5.67051e-8, S, rtol=1e-12)
    K4, L4 = f[2] + f[3] + f[4] + f[5], f[8] + f[9] + f[10] + f[11]
    S_ref = 0.7 * (f[13] * 0.28 + (f[0] + f[1]) * 0.06 + K4 * 0.22) + 0.95 * ((f[6] + f[7]) * 0.06 + f[14] * 0.28 +
                                                                              L4 * 0.22)
    np.testing.assert_allclose(S, S_ref, rtol=1e-14)


def test_perez_v3_drop_in_shape_and_normalisation():
    from solweig_amd import flux, sky
    lv, clear, bright = flux.Perez_v3(35.0, 160.0, 180.0, 540.0, 172, 1, 2)
    assert lv.shape == (153, 3) and abs(lv[:, 2].sum() - 1) < 1e-12
    ref = sky.perez(35.0, 160.0, 180.0, 540.0, 172)
    np.testing.assert_allclose(lv[:, 2], ref[2], rtol=1e-12)      # no patch-sun cosine leaves [-1, 1] here
    assert clear > 1 and bright > 0


ANISO = os.path.join(os.path.dirname(__file__), 'golden', 'umep_aniso_golden.npz')
KSIDE_CASES = 5
LCYL_CASES = 3


def kside_case(g, c):
    """Arguments of umep.kside (UMEP order, masks as float planes) and UMEP's 7 outputs."""
    p = f'k{c}_'
    radI, radD, radG, az, alt, psi, t, albedo, cyl, ani = g[p + 'scalars']
    sh, veg, vb = (m.astype(np.float64) for m in g[p + 'masks'])
    args = [radI, radD, radG, g[p + 'shadow'], g[p + 'svfS'], g[p + 'svfW'], g[p + 'svfN'], g[p + 'svfE'],
            g[p + 'svfEveg'], g[p + 'svfSveg'], g[p + 'svfWveg'], g[p + 'svfNveg'], az, alt, psi, t, albedo,
            g[p + 'F_sh'], g[p + 'KupE'], g[p + 'KupS'], g[p + 'KupW'], g[p + 'KupN'], int(cyl), g[p + 'lv'],
            int(ani), g[p + 'diffsh'], g[p + 'asvf'], sh, veg, vb]
    return args, g[p + 'out']


def lcyl_case(g, c):
    p = f'l{c}_'
    esky, Ta, Tgwall, ewall, salt, sazi = g[p + 'scalars']
    sh, veg, vb = (m.astype(np.float64) for m in g[p + 'masks'])
    args = [esky, g[p + 'patches'], Ta, Tgwall, ewall, g[p + 'Lup'], sh, veg, vb, salt, sazi, g[p + 'asvf']]
    return args, g[p + 'out']


@pytest.mark.parametrize('case', range(KSIDE_CASES))
def test_kside_all_branches_match_umep(case):
    g = np.load(ANISO)
    args, ref = kside_case(g, case)
    out = np.stack([np.broadcast_to(o, ref.shape[1:]) for o in umep.kside(*args)])
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-10)


def test_kside_cylinder_is_the_general_form():
    g = np.load(ANISO)
    args, _ = kside_case(g, 0)
    (radI, radD, _, shadow, *_svf), rest = args[:12], args[12:]
    az, alt, _, _, albedo, _, KupE, KupS, KupW, KupN, _, lv, _, diffsh, asvf, sh, veg, vb = rest
    a = umep.kside(*args)
    b = umep.kside_cylinder(radI, radD, shadow, az, alt, albedo, KupE, KupS, KupW, KupN, lv, diffsh, asvf, sh, veg, vb)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize('case', range(LCYL_CASES))
def test_lcyl_random_masks_match_umep(case):
    g = np.load(ANISO)
    args, ref = lcyl_case(g, case)
    out = np.stack(umep.lcyl(*args))
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-10)
