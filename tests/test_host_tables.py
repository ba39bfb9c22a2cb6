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


def _window_box(wave, off, k0, ke, rows, cols, hb):
    """CPU restatement of the march's window test geometry (window_below): the raster box spanned
    by the wave's pixels shifted by the window's first and last step, its 16 x 16 cells, and the
    height bound over them (four blocks of one level for boxes of at most 8 x 8 cells, else a
    loop over the cells). None when the box leaves the raster entirely."""
    wr0, wr1, wc0, wc1 = wave
    dr0, dr1, dc0, dc1 = off[k0, 0], off[ke, 0], off[k0, 1], off[ke, 1]
    r0, r1 = max(wr0 + min(dr0, dr1), 0), min(wr1 + max(dr0, dr1), rows - 1)
    c0, c1 = max(wc0 + min(dc0, dc1), 0), min(wc1 + max(dc0, dc1), cols - 1)
    if r0 > r1 or c0 > c1:
        return None
    r0c, r1c, c0c, c1c = r0 >> 4, r1 >> 4, c0 >> 4, c1 >> 4
    h, w = r1c - r0c + 1, c1c - c0c + 1
    if h <= 8 and w <= 8:
        i = 2 if h >= 4 else (1 if h >= 2 else 0)
        j = 2 if w >= 4 else (1 if w >= 2 else 0)
        M = hb[i, j]
        rb, cb = r1c - (1 << i) + 1, c1c - (1 << j) + 1
        hm = max(M[r0c, c0c], M[r0c, cb], M[rb, c0c], M[rb, cb])
    else:
        hm = hb[0, 0, r0c:r1c + 1, c0c:c1c + 1].max()
    return (r0, r1, c0, c1), float(hm)


def test_window_box_bounds_every_sample_of_the_window():
    """For waves of every tile width (a wave is 64 lanes: 64 / W rows of W columns, or 64 columns
    of one row) and windows of 32 and 64 steps along sun rays at many azimuths and low altitudes,
    every in-raster sample of every lane lies inside the window's box, and the box's height bound
    is at least max(a, canopy top) of each of them: a window the test skips holds no sample above
    the wave's lowest threshold."""
    from solweig_amd.static import height_bound
    rng = np.random.default_rng(11)
    rows, cols = 300, 233
    a = np.where(rng.random((rows, cols)) < 0.3, rng.uniform(0, 50, (rows, cols)), 0.0)
    t = np.where(rng.random((rows, cols)) < 0.2, rng.uniform(0, 12, (rows, cols)), 0).astype(np.float32)
    top = np.maximum(a, t.astype(np.float64) + a)
    hb = height_bound(a, t)
    windows = 0
    for az in np.arange(3.0, 360.0, 29.0):
        for alt in (2.0, 6.0, 25.0):
            off, dz, *_ = sky.shadow_ray(az, alt, 1.0, 60.0, (rows, cols))
            L = len(dz)
            for tx in (16, 32, 64, 128):
                ty = 256 // tx
                tr0 = int(rng.integers(0, -(-rows // ty))) * ty
                tc0 = int(rng.integers(0, -(-cols // tx))) * tx
                w64 = int(rng.integers(0, 4)) * 64
                sh = tx.bit_length() - 1
                wr0, wr1 = tr0 + (w64 >> sh), tr0 + ((w64 + 63) >> sh)
                wc0 = tc0 + (w64 & (tx - 1)) if tx > 64 else tc0
                wc1 = wc0 + 63 if tx > 64 else tc0 + tx - 1
                lanes = [(min(r, rows - 1), min(c, cols - 1)) for r in range(wr0, wr1 + 1) for c in range(wc0, wc1 + 1)]
                for k0 in range(16, L, 16):
                    for win in (32, 64):
                        ke = min(k0 + win, L) - 1
                        got = _window_box((wr0, wr1, wc0, wc1), off, k0, ke, rows, cols, hb)
                        windows += 1
                        for r, c in lanes:
                            for k in range(k0, ke + 1):
                                rr, cc = r + off[k, 0], c + off[k, 1]
                                if not (0 <= rr < rows and 0 <= cc < cols):
                                    continue
                                assert got is not None, (az, alt, tx, k0)
                                (r0, r1, c0, c1), hm = got
                                assert r0 <= rr <= r1 and c0 <= cc <= c1, (az, alt, tx, k0, k)
                                assert hm >= top[rr, cc]
    assert windows > 1000


def test_hour_branches_follow_the_piecewise_edges():
    """sky.hour_branches changes exactly where the hour chain's pieces change: the Reindl bins at
    clearness index 0.3 and 0.78 (Kt = radG / (1370 cos(zen) D)), and the CI clamp."""
    import numpy as np
    from solweig_amd import sky
    zen, alt, azi, jday, lat, P = np.radians(40.0), 50.0, 180.0, 172.0, 50.0, 101.3
    _, _, kt1 = sky.clearness(zen, jday, 20.0, 0.5, 1.0, lat, P)           # Kt per W/m2 of radG
    for edge, below, above in ((0.3, 0, 1), (0.78, 1, 2)):
        g = edge / kt1
        lo = sky.hour_branches(alt, azi, zen, jday, 20.0, 50.0, g * (1 - 1e-9), P, lat)
        hi = sky.hour_branches(alt, azi, zen, jday, 20.0, 50.0, g * (1 + 1e-9), P, lat)
        assert (lo[2], hi[2]) == (below, above), edge
    night = sky.hour_branches(-5.0, azi, np.radians(95.0), jday, 20.0, 50.0, 0.0, P, lat, CI=0.9)
    assert night == (0, True)
    low = sky.hour_branches(alt, azi, zen, jday, 20.0, 50.0, 100.0, P, lat)
    high = sky.hour_branches(alt, azi, zen, jday, 20.0, 50.0, 1200.0, P, lat)
    assert low[1] is False and high[1] is True                         # CI clamped at high radG


def test_wall_test_breakpoints_reproduce_the_fp64_sum():
    """sky.breakpoints: x + y < t in FP64 exactly when x < c, for x at, just below and just above
    c and at random, over yi values, +inf padding and the band thresholds."""
    thr = np.array([t for alt in (6.0, 18.0, 42.0, 78.0) for t in sky.sunlit_thresholds(alt)])
    rng = np.random.default_rng(7)
    y = np.concatenate([-rng.exponential(2.0, 40), [0.0, -0.0, -1e-300, np.inf]])
    for t in thr:
        c = sky.breakpoints(y, t)
        np.testing.assert_array_equal(c, sky._breakpoints_bisect(y, np.full_like(y, t)))
        xs = np.concatenate([np.abs(rng.normal(0, 3, 2000)), [0.0, 1e300]])
        for x in xs:
            np.testing.assert_array_equal((x + y) < t, x < c)
        for xc in (c, np.nextafter(c, -np.inf), np.nextafter(c, np.inf)):
            ok = np.isfinite(xc) & (xc >= 0)
            np.testing.assert_array_equal(((xc + y) < t)[ok], (xc < c)[ok])


@pytest.mark.parametrize('altitude,azimuth', [(3.5, 70.0), (21.0, 135.0), (48.0, 200.0), (61.0, 300.0)])
def test_search_tables_give_the_direct_wall_classification(altitude, azimuth):
    """The device's bisection over the breakpoint tables (search_row) finds, per altitude band and
    pixel, the patches with tan(asvf) + yi < sunlit threshold and >= shaded threshold, as the
    direct per-patch FP64 test does."""
    from solweig_amd.runtime import SWS_CSH, SWS_CSUN, SWS_MASK
    f, d = sky.patch_rows(1, 0.8, 20.0, altitude, azimuth, lv=np.full(153, 1 / 153), radD=100.0)
    row = sky.search_row(f, d)
    cs = row[SWS_CSUN:SWS_CSUN + 512].view(np.float64).reshape(8, 32)
    ch = row[SWS_CSH:SWS_CSH + 512].view(np.float64).reshape(8, 32)
    mk = row[SWS_MASK:SWS_MASK + 264].reshape(8, 33)
    counts, steps = [31, 30, 28, 24, 19, 13, 7, 1], [5, 5, 5, 5, 5, 4, 3, 1]
    yi = d[:, PD['YI']]
    rng = np.random.default_rng(int(altitude * 10))
    xs = np.concatenate([np.abs(rng.normal(0, 2, 300)), cs[np.isfinite(cs)], ch[np.isfinite(ch)]])
    lo = 0
    for b, n in enumerate(counts):
        t_sun, t_sh = d[lo, PD['SUN_BELOW']], d[lo, PD['SHADE_FROM']]
        bits = np.left_shift(np.uint32(1), np.arange(n, dtype=np.uint32))
        full = np.uint32(0xffffffff) if n == 32 else np.uint32((1 << n) - 1)
        for x in xs:
            js = jh = 0
            s = 1 << (steps[b] - 1)
            while s > 0:
                js += s if x < cs[b, js + s - 1] else 0
                jh += s if x < ch[b, jh + s - 1] else 0
                s >>= 1
            sun = np.bitwise_or.reduce(np.where(x + yi[lo:lo + n] < t_sun, bits, np.uint32(0)))
            shade = np.bitwise_or.reduce(np.where(x + yi[lo:lo + n] >= t_sh, bits, np.uint32(0)))
            assert mk[b, js] == sun, (b, x)
            assert (full & ~mk[b, jh]) == shade, (b, x)
        lo += n


def test_day_search_rows_equal_the_single_hour_rows():
    """build_day fills the breakpoints of all daytime hours in one vectorised bisection; each row
    equals the row search_row builds for its hour alone."""
    d = synthetic.small_domain(seed=3, size=24)
    s = psun.met_sun(d.met, d.location, d.utc)
    tab = sky.build_day(d.met, s, d.latitude, d.dsm.shape, 1.0, 30.0)
    day = np.flatnonzero(tab.hours[:, H['DAY']] != 0)
    assert len(day) > 5
    for i in day:
        np.testing.assert_array_equal(tab.ptab_s[i], sky.search_row(tab.ptab_f[i], tab.ptab_d[i].T))


def test_diffuse_table_holds_each_hour_to_27_bits():
    """sky.diffuse_table (the input of sw_day_diffuse): per daytime hour the luminance row and the
    lumChi cos omega row of its patches (sky positions k, vegetation-only positions 160 + k times
    the hour's vegetation diffsh in FP32), each as integers q on a power-of-two scale, |q| <= 2^27,
    within half a unit of the value, in int8 digits of base 128; night rows carry the scale -1 and
    a row with a non-finite value the scale NaN."""
    from solweig_amd import runtime as rt
    d = synthetic.small_domain(seed=5, size=24, latitude=40.0)
    s = psun.met_sun(d.met, d.location, d.utc)
    t = sky.build_day(d.met[:24], {k: v[:24] for k, v in s.items()}, d.latitude, d.dsm.shape, d.scale, 30.0)
    rows = rt.diff_rows(24)
    assert rows == 48 and t.diffuse.nbytes == rt.diff_table_bytes(24) == rows * (4 * 320 + 8)
    nd = 4 * rows * 320
    digits, scale = sky.diffuse_digits(t.diffuse, 24)
    digits = digits.astype(np.int64)
    # the fragment order of the device: block 1, step 2, digit 3, lane 5 (c 5, g 0) = row 21,
    # positions 128..143 of digit 3
    frag = t.diffuse[:nd].view(np.int8).reshape(3, 5, 4, 64, 16)
    assert np.array_equal(frag[1, 2, 3, 5], digits[3, 21, 128:144])
    assert np.array_equal(frag[0, 4, 0, 63], digits[0, 15, 304:320])
    assert digits[:3].min() >= 0 and digits[:3].max() <= 127
    assert digits[3].min() >= -64 and digits[3].max() <= 64
    q = digits[0] + 128 * digits[1] + 128 ** 2 * digits[2] + 128 ** 3 * digits[3]
    n_day = 0
    for h in range(24):
        r = slice(2 * h, 2 * h + 2)
        if not t.hours[h, H['DAY']]:
            assert np.all(scale[r] == -1.0) and not q[r].any()
            continue
        n_day += 1
        f = t.ptab_f[int(t.hours[h, H['PATCH_TABLE']])].astype(np.float64)
        dveg = float(np.float32(t.hours[h, H['DIFF_VEG']]))
        for v, col in ((0, PF['LV']), (1, PF['KD'])):
            want = np.zeros(320)
            want[:153] = f[:, col]
            want[160:313] = dveg * f[:, col]
            sc = scale[2 * h + v]
            assert np.frexp(sc)[0] == 0.5                               # a power of two
            assert np.abs(q[2 * h + v]).max() <= 2 ** 27
            assert np.abs(want).max() <= 2 ** 27 * sc < 2 * np.abs(want).max()   # 27 bits of the top
            assert np.all(np.abs(q[2 * h + v] * sc - want) <= sc / 2)
    assert n_day >= 10
    # a non-finite patch value: the row's sums are NaN for every pixel (0 x NaN in the reference)
    pf = t.ptab_f.copy()
    h = int(np.flatnonzero(t.hours[:, H['DAY']] != 0)[3])
    pf[int(t.hours[h, H['PATCH_TABLE']]), 7, PF['KD']] = np.nan
    bad_digits, bs = sky.diffuse_digits(sky.diffuse_table(t.hours, pf), 24)
    assert np.isnan(bs[2 * h + 1]) and bs[2 * h] == scale[2 * h]
    assert not bad_digits[:, 2 * h + 1].any() and np.array_equal(bad_digits[:, 2 * h], digits[:, 2 * h])
