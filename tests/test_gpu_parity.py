"""Device path against the oracle (GPU). Masks bit-exact; Tmrt/UTCI within 0.05 K."""
import numpy as np
import pytest

from oracle import day as oday, svf as osvf, umep
from golden_inputs import sun_from_golden

pytestmark = pytest.mark.gpu

TOL_K = 0.05  # Tmrt and UTCI tolerance against the FP64 oracle (kelvin)


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


def _domain(size=40, seed=7, **kw):
    from solweig_amd import synthetic
    return synthetic.small_domain(seed=seed, size=size, **kw)


@pytest.mark.parametrize('size,seed', [(40, 7), (33, 8)])
def test_svf_masks_bit_exact(rt, size, seed):
    from solweig_amd.static import PreparedDomain
    d = _domain(size, seed)
    print(f'svf masks: {size}x{size} seed {seed}', flush=True)
    pd = PreparedDomain(d, keep_host=True)
    shm, vgm, vbm = pd.skyview.masks()
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude)
    np.testing.assert_array_equal(shm, p['shmat'])
    np.testing.assert_array_equal(vgm, p['vegshmat'])
    np.testing.assert_array_equal(vbm, p['vbshvegshmat'])
    planes = pd.host['svf_planes']
    t = osvf.ray_tables(size, size, pd.amaxvalue, 1.0)
    v, total = osvf.reduce(p['shmat'], p['vegshmat'], p['vbshvegshmat'], p['vegdem2'], t)
    np.testing.assert_array_equal(planes[:15], v)
    np.testing.assert_array_equal(planes[15], total)


def test_shadow_masks_bit_exact_every_hour(rt):
    from solweig_amd import batched, sun as psun
    d = _domain(40, 9)
    s = psun.met_sun(d.met, d.location, d.utc)
    eng = batched.Engine([d], n_hours=24)
    masks = rt.DeviceArray((eng.domains[0].P,), np.uint8)
    eng.structs[0].masks = masks.ptr
    rt.call('sw_memcpy_h2d', eng.dom_dev.ptr, __import__('ctypes').addressof(eng.structs),
            __import__('ctypes').sizeof(eng.structs), eng.stream.handle)
    eng.load(eng.tables(sun=[s]))
    eng.reset_state()
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude)
    checked = 0
    for h in range(24):
        if s['altitude'][h] <= 0:
            continue
        print(f'shadow masks: hour {h} alt {s["altitude"][h]:.2f}', flush=True)
        eng.hour(h)
        m = masks.to_host(eng.stream).reshape(d.shape)
        vegsh, sh, vb, wallsh, wallsun, *_ = umep.shadows(
            p['dsm'], p['vegdem'], p['vegdem2'], s['azimuth'][h], s['altitude'][h], 1.0, p['amaxvalue'],
            p['bush'], p['walls'], p['dirwalls'] * np.pi / 180.)
        np.testing.assert_array_equal((m & 1) != 0, sh == 1)
        np.testing.assert_array_equal((m & 2) != 0, vegsh == 1)
        np.testing.assert_array_equal((m & 4) != 0, vb == 1)
        with np.errstate(divide='ignore', invalid='ignore'):
            sunwall = (wallsun / p['walls'] * p['buildings']) == 1
        np.testing.assert_array_equal((m & 8) != 0, sunwall)
        checked += 1
    assert checked >= 10


@pytest.mark.parametrize('plant', [False, True])
def test_shadow_masks_bit_exact_with_and_without_march_plane(rt, plant):
    """The march reads the canopy plane with -inf for absent tops; a planted tree of 2e-15 m on
    a 16 m roof (its quarter vanishes against the roof, itself does not) makes the static prep
    fall back to forming the tops in the march. Both give the oracle's masks bit for bit."""
    from solweig_amd import batched, sun as psun
    d = _domain(40, 23)
    if plant:
        d.dsm = d.dsm.copy()
        d.trees = d.trees.copy()
        d.dsm[5, 5] = np.float32(16.0)
        d.trees[5, 5] = np.float32(2e-15)
    s = psun.met_sun(d.met, d.location, d.utc)
    eng = batched.Engine([d], n_hours=24)
    assert (eng.domains[0].dev['march_trees'] is None) == plant
    masks = rt.DeviceArray((eng.domains[0].P,), np.uint8)
    eng.structs[0].masks = masks.ptr
    rt.call('sw_memcpy_h2d', eng.dom_dev.ptr, __import__('ctypes').addressof(eng.structs),
            __import__('ctypes').sizeof(eng.structs), eng.stream.handle)
    eng.load(eng.tables(sun=[s]))
    eng.reset_state()
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude)
    checked = 0
    for h in range(24):
        if s['altitude'][h] <= 0 or h % 3:
            continue
        print(f'march plane {not plant}: hour {h} alt {s["altitude"][h]:.2f}', flush=True)
        eng.hour(h)
        m = masks.to_host(eng.stream).reshape(d.shape)
        vegsh, sh, vb, *_ = umep.shadows(
            p['dsm'], p['vegdem'], p['vegdem2'], s['azimuth'][h], s['altitude'][h], 1.0, p['amaxvalue'],
            p['bush'], p['walls'], p['dirwalls'] * np.pi / 180.)
        np.testing.assert_array_equal((m & 1) != 0, sh == 1)
        np.testing.assert_array_equal((m & 2) != 0, vegsh == 1)
        np.testing.assert_array_equal((m & 4) != 0, vb == 1)
        checked += 1
    assert checked >= 3


def _day_errors(tmrt, utci, ref):
    et = max(float(np.max(np.abs(tmrt[h] - ref[h]['Tmrt']))) for h in range(len(ref)))
    g = [np.isfinite(ref[h]['UTCI']) for h in range(len(ref))]
    eu = max(float(np.max(np.abs(utci[h][g[h]] - ref[h]['UTCI'][g[h]]))) for h in range(len(ref)))
    nan_ok = all(np.array_equal(np.isnan(utci[h]), ~g[h]) for h in range(len(ref)))
    return et, eu, nan_ok


@pytest.mark.parametrize('kw', [dict(), dict(landcover=True), dict(maps=True), dict(trees=False)])
def test_day_matches_oracle(rt, kw):
    from solweig_amd import batched, sun as psun
    d = _domain(36, 11, **kw)
    print(f'day vs oracle: {kw}', flush=True)
    s = psun.met_sun(d.met, d.location, d.utc)
    tmrt, utci = batched.run_day(d, sun=s)
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude, landcover=d.landcover,
                     alb_grid=d.alb_grid, emis_grid=d.emis_grid)
    ref = oday.run_day(p, d.met, s, keep=('Tmrt',))
    et, eu, nan_ok = _day_errors(tmrt, utci, ref)
    print(f'  max |dTmrt| {et:.3e} K, max |dUTCI| {eu:.3e} K', flush=True)
    assert et < TOL_K and eu < TOL_K and nan_ok


# a seeded sweep: sizes, latitudes from the equator to 60 N, with and without land cover, maps
# and canopy
SWEEP = [(30 + i, 26 + 3 * i, 7.5 * i, dict(landcover=i % 3 == 1, maps=i % 2 == 0, trees=i % 4 != 3))
         for i in range(9)]


@pytest.mark.parametrize('seed,size,lat,kw', SWEEP)
def test_day_matches_oracle_sweep(rt, seed, size, lat, kw):
    from solweig_amd import batched, sun as psun
    d = _domain(size, seed, latitude=lat, **kw)
    print(f'sweep: seed {seed}, {size}x{size}, latitude {lat}, {kw}', flush=True)
    s = psun.met_sun(d.met, d.location, d.utc)
    tmrt, utci = batched.run_day(d, sun=s)
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude, landcover=d.landcover,
                     alb_grid=d.alb_grid, emis_grid=d.emis_grid)
    ref = oday.run_day(p, d.met, s, keep=('Tmrt',))
    et, eu, nan_ok = _day_errors(tmrt, utci, ref)
    print(f'  max |dTmrt| {et:.3e} K, max |dUTCI| {eu:.3e} K', flush=True)
    assert et < TOL_K and eu < TOL_K and nan_ok


@pytest.mark.parametrize('width', [16, 32])
def test_day_matches_oracle_with_staged_tiles(rt, width):
    """96 x 96: the interior tiles walk the ground view from LDS with the compiled step shifts."""
    from solweig_amd import batched, sun as psun
    d = _domain(96, 12, maps=True)
    print(f'day vs oracle 96x96, tile width {width}', flush=True)
    s = psun.met_sun(d.met, d.location, d.utc)
    eng = batched.Engine([d], n_hours=d.met.shape[0])
    eng.set_tiles(width)
    eng.load(eng.tables(sun=[s]))
    eng.reset_state()
    eng.forward_day()
    tmrt, utci = eng.outputs(0)
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude, landcover=d.landcover,
                     alb_grid=d.alb_grid, emis_grid=d.emis_grid)
    ref = oday.run_day(p, d.met, s, keep=('Tmrt',))
    et, eu, nan_ok = _day_errors(tmrt, utci, ref)
    print(f'  max |dTmrt| {et:.3e} K, max |dUTCI| {eu:.3e} K', flush=True)
    assert et < TOL_K and eu < TOL_K and nan_ok


@pytest.mark.parametrize('prefix', ['dayA_', 'dayB_', 'dayC_'])
def test_day_matches_umep_golden(rt, golden, prefix):
    from solweig_amd import batched, synthetic
    loc = golden[prefix + 'loc']
    d = synthetic.Domain(golden[prefix + 'dsm'], golden[prefix + 'trees'], golden[prefix + 'dem'],
                         golden[prefix + 'walls'], golden[prefix + 'aspect'], golden[prefix + 'met'],
                         float(loc[0]), float(loc[1]), float(loc[2]), float(loc[3]),
                         landcover=golden[prefix + 'landcover'] if prefix + 'landcover' in golden else None,
                         alb_grid=golden[prefix + 'alb_grid'] if prefix + 'alb_grid' in golden else None,
                         emis_grid=golden[prefix + 'emis_grid'] if prefix + 'emis_grid' in golden else None)
    print(f'day vs UMEP golden {prefix}', flush=True)
    tmrt, _ = batched.run_day(d, sun=sun_from_golden(golden, prefix))
    err = float(np.max(np.abs(tmrt - golden[prefix + 'step_Tmrt'])))
    print(f'  max |dTmrt| vs UMEP {err:.3e} K', flush=True)
    assert err < TOL_K


@pytest.mark.parametrize('shift', [0.0, -3.0])
def test_svf_calculator_drop_in_matches_oracle(rt, shift):
    """svf.svf_calculator on caller rasters: heights >= 0 run the windowed sun-ray march, a DSM
    partly below 0 (ground at -3 m) the per-pixel march with the reference's off-raster rule
    (sh = a < 0). Masks and the 15 planes + svftotal equal the oracle's."""
    from solweig_amd import svf as psvf
    d = _domain(34, 19)
    a32 = (d.dsm + np.float32(shift)).astype(np.float32)
    a = a32.astype(np.float64)
    t = d.trees.astype(np.float64)
    top = t + a
    top[top == a] = 0
    bot = t * 0.25 + a
    bot[bot == a] = 0
    bush = np.logical_not(bot * top) * top
    amax = float(max(a.max() - a.min(), t.max()))
    print(f'svf_calculator drop-in, DSM shifted by {shift} (min {a.min():.1f})', flush=True)
    got = psvf.svf_calculator(2, amax, a, top, bot, bush, 1.0)
    ref = osvf.svf_calculator(amax, a, top, bot, bush, 1.0)
    for k in range(15):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=f'plane {k}')
    for k in (15, 16, 17):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=f'mask {k}')
    np.testing.assert_array_equal(got[18], ref[18])
