"""Host-side helpers of the device path (CPU): bit packing, thresholds, tables vs the oracle."""
import numpy as np
import pytest

from solweig_amd import sky, static, svf as psvf, sun as psun, synthetic
from solweig_amd.runtime import H, PD, PF
from oracle import day as oday, svf as osvf, umep


def test_visibility_pack_round_trip():
    rng = np.random.default_rng(0)
    masks = [(rng.random((7, 5, 153)) > 0.5).astype(np.float32) for _ in range(3)]
    words = static.pack_visibility(*masks)
    assert words.shape == (3, 5, 35) and words.dtype == np.uint32
    back = psvf.unpack_visibility(words, (7, 5))
    for a, b in zip(masks, back):
        np.testing.assert_array_equal(a, b)
    # bit k of word w is patch 32 w + k
    one = np.zeros((1, 1, 153), np.float32)
    one[0, 0, 37] = 1
    w = static.pack_visibility(one, one, one)
    assert w[0, 1, 0] == 1 << 5 and w[0, 0, 0] == 0


def test_svf_ray_tables_match_oracle():
    a = psvf.ray_tables(50, 40, 23.5, 1.0)
    b = osvf.ray_tables(50, 40, 23.5, 1.0)
    np.testing.assert_array_equal(a['lengths'], b['lengths'])
    cap = a['capacity']
    np.testing.assert_array_equal(a['offsets'], b['offsets'][:, :cap])
    np.testing.assert_array_equal(a['heights'], b['heights'][:, :cap])
    np.testing.assert_array_equal(a['weights'], b['weights'])
    np.testing.assert_array_equal(a['sectors'], b['sectors'])


@pytest.mark.parametrize('alt', [6.0, 18.0, 42.0, 78.0, 90.0, 6.000000000000001])
def test_sunlit_thresholds_split_numpy_arctan(alt):
    lo, hi = sky.sunlit_thresholds(alt)
    assert lo <= hi
    f = lambda x: np.arctan(np.full(16, x)) * (180 / np.pi)
    if np.isfinite(lo) and lo > 0:
        below = np.nextafter(lo, 0)
        assert f(below)[0] < alt and not (f(lo)[0] < alt)
    if np.isfinite(hi) and hi > 0:
        assert f(hi)[0] > alt and not (f(np.nextafter(hi, 0))[0] > alt)


def test_sun_ray_matches_umep_loop():
    from oracle import brute
    for az, al in ((135.0, 30.0), (250.0, 3.0), (10.0, 55.0), (300.0, 8.0), (90.0, 20.0), (45.0, 45.0)):
        off, dz, *_ = sky.shadow_ray(az, al, 1.0, 25.0, (60, 70))
        steps = brute.ray(az, al, 1.0, 25.0, (60, 70))
        assert len(steps) == len(dz)
        np.testing.assert_array_equal(off, np.array([(s[0], s[1]) for s in steps]))
        np.testing.assert_array_equal(dz, np.array([s[2] for s in steps]))


def test_day_tables_follow_the_oracle_scalars():
    d = synthetic.small_domain(seed=3, size=24)
    s = psun.met_sun(d.met, d.location, d.utc)
    t = sky.build_day(d.met, s, d.latitude, d.shape, 1.0, 20.0)
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude)
    state = dict(CI=1., firstdaytime=1., timeadd=0., timestepdec=1 / 24, Tgmap1=0., Tgmap1E=0.,
                 Tgmap1S=0., Tgmap1W=0., Tgmap1N=0., TgOut1=0.)
    for i in range(24):
        hour = dict(altitude=s['altitude'][i], azimuth=s['azimuth'][i], zen=s['zen'][i], jday=s['jday'][i],
                    dectime=s['dectime'][i], altmax=s['altmax'][i], Ta=d.met[i, 11], RH=d.met[i, 10],
                    radG=d.met[i, 14], P=d.met[i, 12], psi=0.03)
        if s['dectime'][i] % 1 == 0:
            state['CI'] = 1.
        out, state = umep.step(p, hour, state)
        m = t.meta[i]
        assert m['CI'] == state['CI']
        assert m['esky'] == out['esky']
        if s['altitude'][i] > 0:
            assert m['radI'] == out['radI'] and m['radD'] == out['radD']
            assert m['tgwall'] == out['Tgwall']


def test_march_canopy_plane():
    """-inf exactly where the FP64 canopy top equals the DSM; None when a quarter canopy vanishes
    against the DSM while the canopy itself does not."""
    from solweig_amd.static import march_trees
    a = np.array([[0.0, 10.0, 16.0, 3.5]])
    t = np.array([[0.0, 4.0, 0.0, 1e-30]], dtype=np.float32)
    m = march_trees(a, t)
    assert m.dtype == np.float32
    assert np.isneginf(m[0, 0]) and m[0, 1] == np.float32(4.0) and np.isneginf(m[0, 2]) and np.isneginf(m[0, 3])
    t2 = np.array([[0.0, 4.0, 2e-15, 0.0]], dtype=np.float32)
    assert march_trees(a, t2) is None


def test_height_bound_block_maxima():
    """Level (0, 0) bounds every sample of its 16 x 16 cell from above; the four-block query of
    the march's window test equals the maximum over any box of at most 8 x 8 cells."""
    from solweig_amd.static import height_bound
    rng = np.random.default_rng(5)
    rows, cols = 150, 117
    a = rng.uniform(0, 40, (rows, cols))
    t = np.where(rng.random((rows, cols)) < 0.2, rng.uniform(0, 12, (rows, cols)), 0).astype(np.float32)
    hb = height_bound(a, t)
    R, C = -(-rows // 16), -(-cols // 16)
    assert hb.shape == (3, 3, R, C) and hb.dtype == np.float32
    top = np.maximum(a, t.astype(np.float64) + a)
    for r in range(R):
        for c in range(C):
            assert float(hb[0, 0, r, c]) >= top[16 * r:16 * r + 16, 16 * c:16 * c + 16].max()
    for _ in range(300):
        r0, c0 = rng.integers(0, R), rng.integers(0, C)
        r1, c1 = min(R - 1, r0 + rng.integers(0, 8)), min(C - 1, c0 + rng.integers(0, 8))
        h, w = r1 - r0 + 1, c1 - c0 + 1
        i = 2 if h >= 4 else (1 if h >= 2 else 0)
        j = 2 if w >= 4 else (1 if w >= 2 else 0)
        M = hb[i, j]
        rb, cb = r1 - (1 << i) + 1, c1 - (1 << j) + 1
        q = max(M[r0, c0], M[r0, cb], M[rb, c0], M[rb, cb])
        assert q == hb[0, 0, r0:r1 + 1, c0:c1 + 1].max()
