"""Solar position on the device (sw_sun_position, FP64 NREL SPA) against UMEP's own values (golden
vectors written by umep_solweig.Solweig_2015a_metdata_noload) and against the host restatement
(sun.met_sun, pinned to the same vectors) over leap years, year boundaries, sub-hourly steps,
polar day and both hemispheres."""
import datetime

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ('altitude', 'azimuth', 'zen', 'jday', 'dectime', 'altmax')
# device against host: the sidereal-time term is about 3e6 degrees before its folding, so its
# FP64 rounding alone moves the angles by ~5e-10 degrees; 1e-8 degrees is 36 micro-arcseconds
ATOL = {'altitude': 1e-8, 'azimuth': 1e-8, 'zen': 2e-10, 'jday': 0.0, 'dectime': 0.0, 'altmax': 1e-8}


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


def _forcing(start, n, step_min):
    rows = []
    for i in range(n):
        t = start + datetime.timedelta(minutes=step_min * i)
        rows.append((t.year, t.timetuple().tm_yday, t.hour, t.minute))
    met = np.full((n, 24), -999.0)
    met[:, :4] = rows
    return met


@pytest.mark.parametrize('k', range(5))
def test_device_sun_matches_umep(rt, golden, k):
    from solweig_amd.sun import met_sun_device
    met = np.full((24, 24), -999.0)
    met[:, :4] = golden[f'met{k}_in']
    lat, lon, utc, alt = golden[f'met{k}_loc']
    got = met_sun_device(met, {'latitude': lat, 'longitude': lon, 'altitude': alt}, utc)
    ref = golden[f'met{k}_out']
    for i, key in enumerate(KEYS):
        np.testing.assert_allclose(got[key], ref[i], rtol=0, atol=1e-9, err_msg=key)


CASES = [
    # (start, rows, step minutes, chunk rows, latitude, longitude, utc)
    (datetime.datetime(2024, 1, 1), 24 * 30, 60, 24, 48.85, 2.35, 1.0),          # leap year, first row half a step into 2023
    (datetime.datetime(2023, 12, 30), 24 * 3, 60, 0, 40.7, -74.0, -5.0),         # one chunk across the new year
    (datetime.datetime(2025, 3, 20), 144, 10, 144, -33.87, 151.21, 10.0),       # 10-minute steps, southern hemisphere
    (datetime.datetime(2025, 6, 21), 48, 60, 24, 78.2, 15.6, 1.0),              # polar day: the sun never sets
    (datetime.datetime(2025, 12, 21), 24, 60, 24, -77.8, 166.7, 12.0),          # polar day in the south
    (datetime.datetime(2025, 9, 1), 24 * 2, 30, 48, 1.3, 103.8, 8.0),           # equatorial, half-hour steps
]


@pytest.mark.parametrize('case', range(len(CASES)))
def test_device_sun_matches_host(rt, case):
    from solweig_amd.sun import met_sun, met_sun_device
    start, n, step, chunk, lat, lon, utc = CASES[case]
    met = _forcing(start, n, step)
    loc = {'latitude': lat, 'longitude': lon, 'altitude': 3.0}
    print(f'sun case {case}: {start:%Y-%m-%d} {n} rows of {step} min, chunks of {chunk}, lat {lat}', flush=True)
    got = met_sun_device(met, loc, utc, chunk_rows=chunk)
    c = chunk or n
    parts = [met_sun(met[i:i + c], loc, utc) for i in range(0, n, c)]
    ref = {key: np.concatenate([p[key] for p in parts]) for key in KEYS}
    for key in KEYS:
        err = float(np.max(np.abs(got[key] - ref[key])))
        print(f'  {key}: max |device - host| {err:.2e}', flush=True)
        assert err <= ATOL[key], key
