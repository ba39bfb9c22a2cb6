"""Batched forms on the GPU: many uneven domains in one launch per pass, and the day loop
replayed from a captured hipGraph. Both must give the same bits as the single-domain eager run,
and every domain stays within 0.05 K of the FP64 oracle (the per-domain max and RMS error of the
FP32 device path against the FP64 oracle is printed)."""
import numpy as np
import pytest

from oracle import day as oday

pytestmark = pytest.mark.gpu

TOL_K = 0.05


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


def _oracle_day(d, s):
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude, landcover=d.landcover,
                     alb_grid=d.alb_grid, emis_grid=d.emis_grid)
    ref = oday.run_day(p, d.met, s, keep=('Tmrt',))
    return np.stack([r['Tmrt'] for r in ref]), np.stack([r['UTCI'] for r in ref])


def test_uneven_domains_in_one_launch(rt):
    """Six domains of sides 21..67 px at latitudes 0-60 N (the C4 generator at small sides)."""
    from solweig_amd import batched, synthetic, sun as psun
    doms = [synthetic.small_domain(seed=40 + i, size=n, latitude=lat, maps=(i % 2 == 1))
            for i, (n, lat) in enumerate([(21, 3.0), (67, 58.0), (33, 24.5), (48, 41.0), (26, 12.0),
                                          (55, 50.0)])]
    suns = [psun.met_sun(d.met, d.location, d.utc) for d in doms]
    eng = batched.Engine(doms, n_hours=24)
    print(f'uneven batch: sides {[d.shape[0] for d in doms]}, {eng.total_blocks} tiles', flush=True)
    eng.load(eng.tables(sun=suns))
    eng.reset_state()
    eng.forward_day()
    for k, (d, s) in enumerate(zip(doms, suns)):
        tmrt, utci = eng.outputs(k)
        t1, u1 = batched.run_day(d, sun=s)
        assert np.array_equal(tmrt, t1, equal_nan=True) and np.array_equal(utci, u1, equal_nan=True), \
            f'domain {k}: batched result differs from its single-domain run'
        rt_, ru = _oracle_day(d, s)
        dt = np.abs(tmrt - rt_)
        g = np.isfinite(ru)
        du = np.abs(utci[g] - ru[g])
        print(f'  domain {k} ({d.shape[0]}x{d.shape[1]}, lat {d.latitude:.1f}): Tmrt max {dt.max():.2e} '
              f'rms {np.sqrt(np.mean(dt ** 2)):.2e} K; UTCI max {du.max():.2e} rms '
              f'{np.sqrt(np.mean(du ** 2)):.2e} K', flush=True)
        assert dt.max() < TOL_K and du.max() < TOL_K
        assert np.array_equal(np.isnan(utci), ~g)


def test_graph_replays_match_eager(rt):
    """Capture the day once, replay it for three different days (tables refreshed in place),
    with eager runs in between; every replay equals the eager run of the same day bit for bit."""
    from solweig_amd import batched, synthetic
    d = synthetic.make_c1(seed=12, size=64, days=3)
    eng = batched.Engine([d], n_hours=24)
    for day in (0, 1, 2, 1):
        print(f'graph replay: day {day}', flush=True)
        eng.load(eng.tables(day_index=day))
        eng.reset_state()
        eng.forward_day(use_graph=False)
        t0, u0 = eng.outputs(0)
        eng.reset_state()
        eng.forward_day(use_graph=True)
        t1, u1 = eng.outputs(0)
        assert np.array_equal(t0, t1, equal_nan=True) and np.array_equal(u0, u1, equal_nan=True)
    assert eng.graph is not None


def test_tile_widths_give_identical_days(rt):
    """Every tile width the autotuner may pick gives the same bits (37 x 53 raster: partial tiles)."""
    from solweig_amd import batched, synthetic
    d = synthetic.small_domain(seed=61, size=53, maps=True)
    d.dsm, d.trees, d.dem = d.dsm[:37], d.trees[:37], d.dem[:37]
    d.walls, d.aspect = d.walls[:37], d.aspect[:37]
    d.alb_grid, d.emis_grid = d.alb_grid[:37], d.emis_grid[:37]
    eng = batched.Engine([d], n_hours=24)
    tabs = eng.tables()
    ref = None
    for w in rt.TILE_CHOICES:
        print(f'tile width {w}', flush=True)
        eng.set_tiles(w)
        eng.load(tabs)
        eng.reset_state()
        eng.forward_day()
        out = eng.outputs(0)
        if ref is None:
            ref = out
        else:
            assert np.array_equal(out[0], ref[0], equal_nan=True) and np.array_equal(out[1], ref[1], equal_nan=True), w


def _staged_tiles(rows, cols, width, second):
    """Tiles whose ground-view halo lies inside the raster (the pixel pass stages them in LDS)."""
    m, height = second - 1, 256 // width
    if width > 32:
        return 0
    return sum(1 for r0 in range(0, rows, height) for c0 in range(0, cols, width)
               if r0 >= m and c0 >= m and r0 + height + m <= rows and c0 + width + m <= cols)


@pytest.mark.parametrize('scale', [1.0, 0.5])
def test_tile_widths_identical_with_staged_tiles(rt, scale):
    """96 x 104 raster: at widths 16 and 32 interior tiles are staged in LDS (at scale 1 with the
    compiled step shifts, at scale 0.5 with the offsets table), at 64 and 128 none are; every
    width gives the same bits."""
    from solweig_amd import batched, synthetic
    from solweig_amd.sky import gvf_geometry
    d = synthetic.small_domain(seed=62, size=104, maps=True)
    d.dsm, d.trees, d.dem = d.dsm[:96], d.trees[:96], d.dem[:96]
    d.walls, d.aspect = d.walls[:96], d.aspect[:96]
    d.alb_grid, d.emis_grid = d.alb_grid[:96], d.emis_grid[:96]
    d.scale = scale
    second = gvf_geometry(scale)[3]
    assert _staged_tiles(96, 104, 16, second) > 0 and _staged_tiles(96, 104, 32, second) > 0
    eng = batched.Engine([d], n_hours=24)
    tabs = eng.tables()
    ref = None
    for w in rt.TILE_CHOICES:
        print(f'scale {scale} tile width {w}: {_staged_tiles(96, 104, w, second)} staged tiles', flush=True)
        eng.set_tiles(w)
        eng.load(tabs)
        eng.reset_state()
        eng.forward_day()
        out = eng.outputs(0)
        if ref is None:
            ref = out
        else:
            assert np.array_equal(out[0], ref[0], equal_nan=True) and np.array_equal(out[1], ref[1], equal_nan=True), w


def test_autotuner_picks_a_measured_width(rt, tmp_path, monkeypatch):
    from solweig_amd import autotune, batched, synthetic
    monkeypatch.setattr(autotune, 'CACHE', str(tmp_path / 'tiles.json'))
    d = synthetic.make_c1(seed=13, size=96, days=1)
    eng = batched.Engine([d], n_hours=24)
    res = autotune.tune(eng, repeats=2)
    print(f'autotune: {res}', flush=True)
    assert res['tile_cols'] in rt.TILE_CHOICES and eng.domains[0].tile_cols == res['tile_cols']
    assert set(res['ms']) == set(rt.TILE_CHOICES)
    again = autotune.tune(eng)
    assert again['cached'] and again['tile_cols'] == res['tile_cols']


@pytest.mark.parametrize('tile', [16, 32])
def test_resident_day_equals_hourly_launches(rt, tile):
    """sw_day_forward (all shadow passes, then one pixel pass with each tile running through the
    hours with its lag state in registers) gives the bits of the two-launches-per-hour loop, over
    uneven domains, for a day split at hour 10 (lag state stored and reloaded between the two
    launches) and for graph replays."""
    from solweig_amd import batched, synthetic, sun as psun
    doms = [synthetic.small_domain(seed=60 + i, size=n, latitude=lat, maps=(i % 2 == 0))
            for i, (n, lat) in enumerate([(70, 8.0), (37, 47.0), (52, 33.0)])]
    suns = [psun.met_sun(d.met, d.location, d.utc) for d in doms]
    runs = {}
    for resident in (False, True):
        eng = batched.Engine(doms, n_hours=24, resident=resident)
        eng.set_tiles(tile)
        eng.load(eng.tables(sun=suns))
        print(f'tile {tile}, resident {resident}: whole day', flush=True)
        eng.reset_state()
        eng.forward_day()
        runs[resident, 'day'] = [eng.outputs(k) for k in range(len(doms))]
        print(f'tile {tile}, resident {resident}: hours 0-10 and 10-24', flush=True)
        eng.reset_state()
        eng.run_hours(0, 10)
        eng.run_hours(10, 24)
        runs[resident, 'split'] = [eng.outputs(k) for k in range(len(doms))]
        if resident:
            for rep in range(2):
                print(f'tile {tile}, resident: graph replay {rep}', flush=True)
                eng.reset_state()
                eng.forward_day(use_graph=True)
                runs[resident, f'graph{rep}'] = [eng.outputs(k) for k in range(len(doms))]
    ref = runs[False, 'day']
    for key, got in runs.items():
        for k, ((t, u), (t0, u0)) in enumerate(zip(got, ref)):
            assert np.array_equal(t, t0) and np.array_equal(u, u0, equal_nan=True), (key, k)


@pytest.mark.parametrize('use_graph', [False, True])
def test_pipelined_days_equal_single_days(rt, use_graph):
    """Engine.run_days (each day's all-hours shadow pass on a side stream beside the previous day's
    pixel pass, days alternating between two table and record sets) gives every day the bits of the
    same day run on its own, over uneven domains; a second run on the same engine, in another day
    order, starts again from the first set."""
    from solweig_amd import autotune, batched, synthetic
    doms = [synthetic.make_c1(seed=70 + i, size=n, days=4) for i, n in enumerate((48, 33, 61))]
    alone = batched.Engine(doms, n_hours=24)
    tabs = [alone.tables(day_index=day) for day in range(4)]
    ref = []
    for day in range(4):
        alone.load(tabs[day])
        alone.reset_state()
        alone.forward_day()
        ref.append([alone.outputs(k) for k in range(len(doms))])
    eng = batched.Engine(doms, n_hours=24, pipelined=True)
    for order in ([0, 1, 2, 3], [3, 1, 2]):
        print(f'pipelined days {order}, graph {use_graph}', flush=True)
        got = []
        n = eng.run_days((tabs[day] for day in order), use_graph=use_graph,
                         after_day=lambda i: got.append([eng.outputs(k) for k in range(len(doms))]))
        assert n == len(order) and not eng.day_open
        for day, outs in zip(order, got):
            for k, ((t, u), (t0, u0)) in enumerate(zip(outs, ref[day])):
                assert np.array_equal(t, t0) and np.array_equal(u, u0, equal_nan=True), (order, day, k)
    # while a run is open its days own both table sets: loads and single-day launches are refused
    eng.day_start(tabs[0])
    with pytest.raises(RuntimeError, match='pipelined days is open'):
        eng.load(tabs[1])
    with pytest.raises(RuntimeError, match='pipelined days is open'):
        eng.forward_day()
    eng.day_next(None)
    assert not eng.day_open
    # the launch-form tuner runs all three forms and leaves one set
    res = autotune.tune_launch_form(eng, tabs, days=3, rounds=1, use_graph=use_graph)
    assert res['form'] in autotune.FORMS and set(res['ms']) == set(autotune.FORMS)
    assert eng.pipelined == (res['form'] == 'pipelined')
