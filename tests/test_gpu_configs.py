"""The benchmark configurations on the device against the FP64 oracle (GPU).

* C0 in full (200 x 200, 12 boxes, 24 h) under both readings of its land cover (all paved, and
  the default maps of series.py:307-315);
* the C1 generator (35 % boxes, 15 % canopy, cloudy forcing) at 256 x 256 for 24 h;
* two leaf-off days (psi = 0.5: January at 35 N, November at 10 N);
* C2 at its real size (2000 x 2000, one day): tile widths 16 and 32 and hipGraph replays give
  identical bits, and the masks of a 15-20 degree hour equal the oracle's;
* long rays: with the sun at 3-5 degrees over a 512 x 512 C1 raster (rays of ~500 steps, so the
  skipped windows, both height-bound queries and the shadowed-wall threshold drop all act), the
  masks equal the oracle's at tile widths 16 and 128.

Masks (sky view and every daytime hour's shadow byte) must be bit-exact; Tmrt and UTCI within
0.05 K. The oracle legs run in spawned CPU worker processes (tests/oracle_jobs.py) while the
device work proceeds.
"""
import concurrent.futures as cf
import multiprocessing as mp

import numpy as np
import pytest

import oracle_jobs as oj

pytestmark = pytest.mark.gpu

TOL_K = 0.05          # Tmrt and UTCI tolerance against the FP64 oracle (kelvin)
DAY_JOBS = ('c1_256', 'c0_paved', 'c0_default', 'leafoff_jan', 'leafoff_nov')


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


@pytest.fixture(scope='module')
def jobs(rt):
    pool = cf.ProcessPoolExecutor(max_workers=7, mp_context=mp.get_context('spawn'))
    fut = {n: pool.submit(oj.day_job, n) for n in DAY_JOBS}
    fut['c2_2000'] = pool.submit(oj.shadow_job, 'c2_2000')
    fut['long_512'] = pool.submit(oj.shadow_job, 'long_512', list(range(5, 9)))
    yield fut
    pool.shutdown(wait=False, cancel_futures=True)


def _engine(d, s, tile=None):
    from solweig_amd import batched
    eng = batched.Engine([d], n_hours=d.met.shape[0])
    if tile:
        eng.set_tiles(tile)
    eng.load(eng.tables(sun=[s]))
    eng.reset_state()
    return eng


def _step_hours(eng, s, hours, keep):
    """Run ``hours`` one by one (both passes) and return the shadow pass's mask byte of the
    daytime hours in ``keep``."""
    out = {}
    for h in hours:
        eng.hour(h)
        if h in keep and s['altitude'][h] > 0:
            out[h] = eng.hour_masks(0)
    return out


def _errors(tmrt, utci, ref_tmrt, ref_utci):
    et = float(np.max(np.abs(tmrt - ref_tmrt)))
    g = np.isfinite(ref_utci)
    eu = float(np.max(np.abs(utci[g] - ref_utci[g])))
    return et, eu, bool(np.array_equal(np.isnan(utci), ~g))


@pytest.mark.timeout(300)
@pytest.mark.parametrize('name', DAY_JOBS)
def test_config_day_matches_oracle(rt, jobs, name):
    d = oj.domain(name)
    s = oj.sun(name, d)
    print(f'config day {name}: {d.shape[0]}x{d.shape[1]}, DOY {int(d.met[0, 1])}, latitude {d.latitude}, '
          f'land cover {"paved" if d.landcover is not None else "default"}', flush=True)
    eng = _engine(d, s)
    masks = _step_hours(eng, s, range(eng.n_hours), range(eng.n_hours))
    tmrt, utci = eng.outputs(0)
    sky = eng.domains[0].skyview.masks()
    ref = jobs[name].result(timeout=280)
    for got, exp, what in zip(sky, ref['svf'], ('shmat', 'vegmat', 'vbmat')):
        np.testing.assert_array_equal(np.packbits(got != 0, axis=-1), exp, err_msg=what)
    assert sorted(masks) == sorted(ref['masks'])
    for h in masks:
        np.testing.assert_array_equal(masks[h], ref['masks'][h], err_msg=f'hour {h} mask byte')
    et, eu, nan_ok = _errors(tmrt, utci, ref['tmrt'], ref['utci'])
    print(f'  {len(masks)} daytime hours, masks identical; max |dTmrt| {et:.3e} K, max |dUTCI| {eu:.3e} K, '
          f'psi {sorted(set(ref["psi"].tolist()))}', flush=True)
    if name.startswith('leafoff'):
        assert (ref['psi'] == 0.5).all()
    assert et < TOL_K and eu < TOL_K and nan_ok


@pytest.mark.timeout(300)
def test_c2_2000_day_properties(rt, jobs):
    d = oj.domain('c2_2000')
    s = oj.sun('c2_2000', d)
    check = int(np.flatnonzero((s['altitude'] >= 15.0) & (s['altitude'] <= 20.0))[0])
    print(f'c2 2000x2000: tile 16 eager, masks at hour {check} (altitude {s["altitude"][check]:.1f})', flush=True)
    eng = _engine(d, s, tile=16)
    masks = _step_hours(eng, s, range(eng.n_hours), {check})
    tmrt, utci = eng.outputs(0)
    assert np.isfinite(tmrt).all() and 0 < np.isnan(utci).mean() < 1
    print('c2 2000x2000: tile 32 eager', flush=True)
    eng.set_tiles(32)
    eng.reset_state()
    eng.forward_day()
    t32, u32 = eng.outputs(0)
    assert np.array_equal(t32, tmrt) and np.array_equal(u32, utci, equal_nan=True)
    eng.set_tiles(16)
    for rep in range(3):
        print(f'c2 2000x2000: graph replay {rep}', flush=True)
        eng.reset_state()
        eng.forward_day(use_graph=True)
        tg, ug = eng.outputs(0)
        assert np.array_equal(tg, tmrt) and np.array_equal(ug, utci, equal_nan=True), rep
    ref = jobs['c2_2000'].result(timeout=280)
    assert list(ref) == [check]
    np.testing.assert_array_equal(masks[check], ref[check])
    print(f'  widths 16/32 and 3 graph replays identical; hour {check} masks identical '
          f'({100 * (1 - (masks[check] & 1).mean()):.1f} % building shadow)', flush=True)


@pytest.mark.timeout(300)
@pytest.mark.parametrize('tile', [16, 128])
def test_long_ray_masks_bit_exact(rt, jobs, tile):
    d = oj.domain('long_512')
    s = oj.sun('long_512', d)
    print(f'long rays 512x512, tile width {tile}: suns {oj.LOW_SUNS}', flush=True)
    eng = _engine(d, s, tile=tile)
    masks = _step_hours(eng, s, range(9), range(5, 9))
    ref = jobs['long_512'].result(timeout=280)
    for h in range(5, 9):
        np.testing.assert_array_equal(masks[h], ref[h], err_msg=f'hour {h} (sun {oj.LOW_SUNS[h - 5]})')
    shaded = [float(1 - (masks[h] & 1).mean()) for h in range(5, 9)]
    print(f'  masks identical; building-shadow fractions {np.round(shaded, 3).tolist()}', flush=True)


@pytest.mark.timeout(300)
@pytest.mark.parametrize('name', ['leafoff_jan', 'c0_paved'])
def test_fp64_series_day_matches_oracle(rt, jobs, name):
    """The dtype=float64 series day (the reference's step decomposition in FP64 on the device,
    outputs rounded to float32 as the reference's output path writes them) against the FP64 oracle
    day: Tmrt within half a float32 ulp of the oracle plus 1e-6 K, UTCI (from the float32 Tmrt)
    within 1e-3 K, the same NaN pattern."""
    from solweig_amd import series
    d = oj.domain(name)
    s = oj.sun(name, d)
    print(f'fp64 series day {name}: {d.shape[0]}x{d.shape[1]}', flush=True)
    ctx = series.stepwise_context(d)
    tmrt, utci, _ = series.stepwise_day(*ctx, d.met, s)
    ref = jobs[name].result(timeout=280)
    excess = np.abs(tmrt.astype(np.float64) - ref['tmrt']) - 0.5 * np.spacing(np.abs(tmrt)).astype(np.float64)
    g = np.isfinite(ref['utci'])
    eu = float(np.max(np.abs(utci[g] - ref['utci'][g])))
    print(f'  Tmrt beyond float32 rounding {float(excess.max()):.2e} K; max |dUTCI| {eu:.2e} K', flush=True)
    assert float(excess.max()) < 1e-6
    assert eu < 1e-3 and np.array_equal(np.isnan(utci), ~g)
