"""The wall-height / wall-aspect oracle pinned by closed forms (boxes on flat ground), and the
device template tables against the oracle's."""
import numpy as np
import pytest

from oracle import walls as ow


def _box(rows=40, cols=44, h=10.0, r0=12, r1=25, c0=15, c1=31):
    a = np.zeros((rows, cols))
    a[r0:r1, c0:c1] = h
    return a


def test_heights_ring_of_a_box():
    a = _box()
    w = ow.wall_heights(a, 3.0)
    ring = np.zeros_like(a, dtype=bool)
    ring[11, 15:31] = ring[25, 15:31] = True
    ring[12:25, 14] = ring[12:25, 31] = True
    assert np.array_equal(w > 0, ring)
    assert np.all(w[ring] == 10.0)
    assert not np.any(ow.wall_heights(_box(h=2.5), 3.0))
    assert np.array_equal(ow.wall_heights(_box(h=3.0), 3.0) > 0, ring)     # threshold inclusive


def test_first_row_quirk():
    a = np.zeros((5, 12))
    a[0, :] = -4.0
    w = ow.wall_heights(a, 3.0)
    assert np.array_equal(np.flatnonzero(w[0]), np.arange(5, 11))          # columns rows..cols-2
    assert np.all(w[0, 5:11] == 4.0) and not np.any(w[1:])


def test_aspect_faces_away_from_the_box():
    a = _box()
    w, asp = ow.dsm_walls(a, 1.0, 3.0)
    west, east = asp[12:25, 14], asp[12:25, 31]
    north, south = asp[11, 15:31], asp[25, 15:31]
    assert np.all(west == 270.0) and np.all(east == 90.0)
    assert np.all((north <= 10.0) | (north >= 350.0))
    assert np.all((south >= 170.0) & (south <= 190.0))
    assert np.all(asp[w == 0] == 0.0)


def test_gradient_fallback_near_the_border():
    # the west wall's top pixels lie within the template radius of the border: the gradient
    # (central differences, +h/2 towards the box) gives 270 degrees, as the templates do below
    a = _box(rows=30, cols=30, r0=1, r1=20, c0=10, c1=20)
    w, asp = ow.dsm_walls(a, 1.0, 3.0)
    assert w[2, 9] > 0 and asp[2, 9] == pytest.approx(270.0, abs=1e-12)
    assert asp[15, 9] == 270.0


def test_small_raster_uses_gradients_only():
    a = _box(rows=12, cols=12, h=6.0, r0=4, r1=8, c0=4, c1=8)
    w, asp = ow.dsm_walls(a, 1.0, 3.0)
    assert np.all(asp[5:7, 3] == pytest.approx(270.0)) and np.all(asp[5:7, 8] == pytest.approx(90.0))
    assert np.all(asp[3, 5:7] == 0.0) and np.all(asp[8, 5:7] == pytest.approx(180.0))


@pytest.mark.parametrize('size', [3, 9, 19])
def test_device_template_tables_match_oracle(size):
    from solweig_amd import walls
    line, ls, side, ss = walls.template_tables(size)
    t = ow.templates(size)
    for k in range(180):
        assert np.array_equal(line[ls[k]:ls[k + 1]], t[k][0])
        assert np.array_equal(side[ss[k, 0]:ss[k, 1]], t[k][1])
        assert np.array_equal(side[ss[k, 1]:ss[k, 2]], t[k][2])
