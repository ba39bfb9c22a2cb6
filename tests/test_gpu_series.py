"""The time-series driver on the GPU: walls from the DSM on the device, every day through the
batched engine, outputs through the asynchronous output path (device staging, side-stream copy,
writer thread). Written days must equal the batched engine's own outputs bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


@pytest.mark.parametrize('slots', [1, 2])
def test_run_timeseries_writes_every_day(rt, tmp_path, slots):
    from solweig_amd import batched, series, synthetic, walls
    d = synthetic.make_c1(seed=21, size=56, days=3)
    for name in ('dsm', 'trees', 'dem'):
        np.save(tmp_path / f'{name}.npy', getattr(d, name))
    mask = np.ones(d.shape, dtype=bool)
    mask[:5] = False
    print(f'run_timeseries: 3 days, {slots} output slots', flush=True)
    res = series.run_timeseries(tmp_path / 'dsm.npy', tmp_path / 'trees.npy', tmp_path / 'dem.npy', d.met,
                                tmp_path / 'out', d.longitude, d.latitude, utc_offset_hours=0.0,
                                output_queue_size=slots, output_valid_mask=mask)
    assert len(res['days']) == 3
    w, a = walls.calculate_dsm_walls(d.dsm.astype(np.float64), 1.0, 3.0)
    dom = synthetic.Domain(d.dsm, d.trees, d.dem, w.astype(np.float32), a.astype(np.float32), d.met, d.latitude,
                           d.longitude, 0.0)
    for k, (date, files) in enumerate(res['days'].items()):
        tmrt = np.load(files['TMRT'][0])
        utci = np.load(files['UTCI'][0])
        t1, u1 = batched.run_day(dom, met=d.met[24 * k:24 * (k + 1)])
        u1 = np.where(mask[None], u1, np.nan)
        t1 = np.where(mask[None], t1, np.nan)          # the valid mask applies to every variable
        assert np.array_equal(tmrt, t1, equal_nan=True), date
        assert np.array_equal(utci, u1, equal_nan=True), date


def test_run_timeseries_no_write_returns_arrays(rt, tmp_path):
    from solweig_amd import series, synthetic
    d = synthetic.make_c1(seed=22, size=40, days=1)
    res = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path, d.longitude, d.latitude,
                                utc_offset_hours=0.0, no_write=True, return_arrays=True)
    (day,) = res['days'].values()
    assert day['UTCI'].shape == (24, 40, 40) and np.isfinite(day['TMRT']).all()
    # the reference's no-write mode keeps nothing: every day maps to an empty dict
    res = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path, d.longitude, d.latitude,
                                utc_offset_hours=0.0, no_write=True)
    assert list(res['days'].values()) == [{}] and not (tmp_path / 'static').exists()


def test_run_timeseries_extra_variables_chunks_and_static(rt, tmp_path):
    """save_kup/kdown/lup/ldown/shadow through the batched pass's extra planes, split into
    10-band files with their tags; they agree with the per-step full-mode planes (plain=True),
    and the static walls, aspect and SVF files are written."""
    import json
    from solweig_amd import series, synthetic
    d = synthetic.make_c1(seed=23, size=48, days=1)
    kw = dict(utc_offset_hours=0.0, save_kup=True, save_kdown=True, save_lup=True, save_ldown=True,
              save_shadow=True, save_svf=True, output_chunk_steps=10)
    print('run_timeseries: extra variables, 10-band chunks, batched and plain', flush=True)
    res = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path / 'b', d.longitude, d.latitude, **kw)
    ref = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path / 'p', d.longitude, d.latitude, plain=True,
                                **kw)
    assert set(res['static']) == {'walls', 'aspect', 'SVF'}
    svf = np.load(res['static']['SVF'])
    assert svf.shape == (48, 48) and 0 < svf.min() and svf.max() <= 1
    (date, files), = res['days'].items()
    assert set(files) == {'UTCI', 'TMRT', 'Kup', 'Kdown', 'Lup', 'Ldown', 'Shadow'}
    for var, paths in files.items():
        assert [p.rsplit('/', 1)[1] for p in paths] == [f'{var}_000{k}.npy' for k in (1, 2, 3)]
        got = np.concatenate([np.load(p) for p in paths])
        exp = np.concatenate([np.load(p) for p in ref['days'][date][var]])
        assert got.shape == (24, 48, 48)
        tags = json.loads(open(paths[2].replace('.npy', '.json')).read())
        assert tags['variable'] == var and len(tags['bands']) == 4
        if var == 'Shadow':
            np.testing.assert_array_equal(got, exp)
        elif var != 'UTCI':
            np.testing.assert_allclose(got, exp, rtol=2e-4, atol=2e-2, err_msg=var)


def test_run_timeseries_fp64(rt, tmp_path):
    """dtype=np.float64 (run.py:28) runs the day through the FP64 step decomposition; it stays
    within the FP32 bar of the batched FP32 series and keeps its NaN pattern and shadow."""
    from solweig_amd import series, synthetic
    d = synthetic.make_c1(seed=24, size=40, days=1)
    kw = dict(utc_offset_hours=0.0, no_write=True, return_arrays=True, save_shadow=True, save_kdown=True)
    print('run_timeseries: float64 vs float32', flush=True)
    r64 = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path, d.longitude, d.latitude,
                                dtype=np.float64, **kw)
    r32 = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path, d.longitude, d.latitude, **kw)
    (a,), (b,) = r64['days'].values(), r32['days'].values()
    assert set(a) == set(b) == {'UTCI', 'TMRT', 'Shadow', 'Kdown'}
    for v in a:
        assert a[v].dtype == np.float32 and np.array_equal(np.isnan(a[v]), np.isnan(b[v])), v
    assert float(np.nanmax(np.abs(a['TMRT'] - b['TMRT']))) < 0.05
    assert float(np.nanmax(np.abs(a['UTCI'] - b['UTCI']))) < 0.05
    np.testing.assert_allclose(a['Shadow'], b['Shadow'], atol=1e-6)
    np.testing.assert_allclose(a['Kdown'], b['Kdown'], rtol=2e-4, atol=2e-2)
    with pytest.raises(ValueError):
        series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path, d.longitude, d.latitude, dtype=np.float16,
                              **kw)
