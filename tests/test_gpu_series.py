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
        assert np.array_equal(tmrt, t1, equal_nan=True), date
        assert np.array_equal(utci, u1, equal_nan=True), date


def test_run_timeseries_no_write_returns_arrays(rt, tmp_path):
    from solweig_amd import series, synthetic
    d = synthetic.make_c1(seed=22, size=40, days=1)
    res = series.run_timeseries(d.dsm, d.trees, d.dem, d.met, tmp_path, d.longitude, d.latitude,
                                utc_offset_hours=0.0, no_write=True)
    (day,) = res['days'].values()
    assert day['UTCI'].shape == (24, 40, 40) and np.isfinite(day['TMRT']).all()
