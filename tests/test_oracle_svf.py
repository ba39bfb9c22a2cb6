"""Pins of the sky-view oracle (CPU): the vectorised visibility march and its ray tables against
an explicit per-pixel loop, against the UMEP-pinned shadow march at the patch centres, and
against closed forms (flat ground, an infinitely tall wall)."""
import numpy as np
import pytest

from oracle import brute, svf as osvf, umep
from solweig_amd import svf as psvf


def _inputs(seed, rows, cols, negative=False):
    """Boxes, canopy blobs and bush pixels formed as the series forms them (series.py:282-287),
    plus forced bush pixels (canopy top > 1 with no trunk zone) and, optionally, pixels below 0."""
    rng = np.random.default_rng(seed)
    a = np.zeros((rows, cols))
    for _ in range(4):
        h, w = rng.integers(2, 6, size=2)
        r, c = rng.integers(0, rows - h), rng.integers(0, cols - w)
        a[r:r + h, c:c + w] = np.maximum(a[r:r + h, c:c + w], rng.uniform(3.0, 12.0))
    t = np.zeros((rows, cols))
    for _ in range(3):
        r, c = rng.uniform(0, rows), rng.uniform(0, cols)
        yy, xx = np.mgrid[0:rows, 0:cols]
        t = np.maximum(t, rng.uniform(3.0, 9.0) * np.exp(-((yy - r) ** 2 + (xx - c) ** 2) / 6.0))
    t[(a > 0) | (t < 1.0)] = 0.0
    if negative:
        a[rng.random(a.shape) < 0.05] = -0.5
    top = t + a
    top[top == a] = 0
    bottom = t * 0.25 + a
    bottom[bottom == a] = 0
    sel = (rng.random(a.shape) < 0.06) & (a >= 0)
    top[sel] = a[sel] + rng.uniform(1.5, 6.0, size=int(sel.sum()))
    bottom[sel] = 0.0
    bush = np.logical_not(bottom * top) * top
    amax = max(a.max() - a.min(), t.max(), top.max() - a.min())
    return a, top, bottom, bush, float(amax)


@pytest.mark.parametrize('seed,rows,cols,scale,negative', [(1, 18, 22, 1.0, False), (2, 20, 16, 0.5, True),
                                                           (3, 16, 16, 2.0, False)])
def test_visibility_matches_per_pixel_loop(seed, rows, cols, scale, negative):
    """All 153 x 3 masks of the shifted-array march equal the per-(pixel, patch) loop bit for bit."""
    a, top, bottom, bush, amax = _inputs(seed, rows, cols, negative)
    assert (bush > 1).any() and (top > 0).any()
    t = osvf.ray_tables(rows, cols, amax, scale)
    got = osvf.visibility(a, top, bottom, bush, t)
    ref = brute.sky_masks(a, top, bottom, bush, scale, amax)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    # the masks are not trivial: some patches hidden, some seen through canopy, some both ways
    assert 0 < got[0].mean() < 1 and 0 < got[1].mean() < 1 and got[2].min() == 0


def test_visibility_equals_umep_shadow_march_at_patch_centres():
    """A patch is seen iff a sun at the patch centre would light the pixel: the three masks equal
    (sh, vegsh, vbshvegsh) of the UMEP-pinned shadow march (shadowingfunction_wallheight_23)."""
    a, top, bottom, bush, amax = _inputs(4, 20, 18)
    t = osvf.ray_tables(20, 18, amax, 1.0)
    shm, vgm, vbm = osvf.visibility(a, top, bottom, bush, t)
    alt, azi = umep.create_patches(2)
    z = np.zeros_like(a)
    for k in range(len(alt)):
        vegsh, sh, vb, *_ = umep.shadows(a, top, bottom, azi[k], alt[k], 1.0, amax, bush, z, z)
        np.testing.assert_array_equal(shm[:, :, k], sh, err_msg=f'patch {k}')
        np.testing.assert_array_equal(vgm[:, :, k], vegsh, err_msg=f'patch {k}')
        np.testing.assert_array_equal(vbm[:, :, k], vb, err_msg=f'patch {k}')


@pytest.mark.parametrize('rows,cols,amax,scale', [(50, 40, 23.5, 1.0), (17, 63, 80.0, 0.5), (64, 9, 5.0, 2.0),
                                                  (12, 12, 0.0, 1.0), (10, 14, -1.0, 1.0)])
def test_ray_tables_are_the_sun_rays_at_patch_centres(rows, cols, amax, scale):
    """Every patch's ray (offsets, heights, length) is the UMEP loop ray of a sun at the patch
    centre, index from 0; the product tables equal the oracle's."""
    t = osvf.ray_tables(rows, cols, amax, scale)
    alt, azi = umep.create_patches(2)
    for k in range(len(alt)):
        steps = brute.ray(azi[k], alt[k], scale, amax, (rows, cols))
        n = int(t['lengths'][k])
        assert n == len(steps), k
        if n:
            np.testing.assert_array_equal(t['offsets'][k, :n], np.array([(s[0], s[1]) for s in steps]))
            np.testing.assert_array_equal(t['heights'][k, :n], np.array([s[2] for s in steps]))
    p = psvf.ray_tables(rows, cols, amax, scale)
    cap = p['capacity']
    np.testing.assert_array_equal(p['lengths'], t['lengths'])
    np.testing.assert_array_equal(p['offsets'], t['offsets'][:, :cap])
    np.testing.assert_array_equal(p['heights'], t['heights'][:, :cap])
    np.testing.assert_array_equal(p['weights'], t['weights'])
    np.testing.assert_array_equal(p['sectors'], t['sectors'])
    np.testing.assert_array_equal(p['band_offsets'], t['band_offsets'])


def test_sector_bits_are_compass_half_circles():
    """Bit d (E, S, W, N) is set iff the patch azimuth lies within 90 degrees (right-open) of the
    compass direction 90 (d + 1) mod 360."""
    alt, azi = umep.create_patches(2)
    bits = osvf.ray_tables(4, 4, 1.0, 1.0)['sectors']
    for d, centre in enumerate((90.0, 180.0, 270.0, 0.0)):
        lo = (centre - 90.0) % 360.0
        inside = ((azi - lo) % 360.0) < 180.0
        np.testing.assert_array_equal((bits >> d) & 1, inside.astype(np.uint8))


def test_half_sky_beside_an_infinitely_tall_wall():
    """A wall along the whole north edge, taller than any ray can climb before reaching it: a
    ground pixel just south of it sees exactly the patches whose azimuth points south of the
    east-west line (90 < azimuth < 270, none lies on it) plus the zenith patch. svf is the band
    weights times those counts; svfN is the zenith patch alone; close to the continuous half sky
    (1 + sin^2 6 deg) / 2, which includes half the 6-degree zenith cap."""
    rows, cols, H = 40, 64, 1000.0
    a = np.zeros((rows, cols))
    a[:10] = H
    z = np.zeros_like(a)
    t = osvf.ray_tables(rows, cols, H, 1.0)
    shm, vgm, vbm = osvf.visibility(a, z, z, z, t)
    v, total = osvf.reduce(shm, vgm, vbm, z, t)
    alt, azi = umep.create_patches(2)
    seen = ((azi > 90) & (azi < 270)) | (alt == 90)
    np.testing.assert_array_equal(shm[10, 32].astype(bool), seen)
    counts = [int(seen[t['band_offsets'][b]:t['band_offsets'][b + 1]].sum()) for b in range(8)]
    expect = 0.0
    for b in range(8):
        expect += t['weights'][b, 0] * counts[b]
    assert v[0][10, 32] == pytest.approx(expect, rel=1e-14)
    assert v[4][10, 32] == t['weights'][7, 1]                     # svfN: the zenith patch only
    assert abs(v[0][10, 32] - 0.5 * (1 + np.sin(np.radians(6.0)) ** 2)) < 0.02
    # the weights integrate the cosine-weighted vault exactly: sum over all patches = 1
    assert float(t['weights'][:, 0] @ np.diff(t['band_offsets'])) == pytest.approx(1.0, abs=1e-12)
