"""Device wall heights and aspect (walls.calculate_dsm_walls) against the oracle: heights and
template-resolved aspects exactly, gradient-fallback aspects to FP64 round-off."""
import numpy as np
import pytest

from oracle import walls as ow

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


def _city(seed, rows, cols, boxes=8, noise=0.0):
    rng = np.random.default_rng(seed)
    a = np.zeros((rows, cols))
    for _ in range(boxes):
        h = int(rng.integers(2, max(3, min(16, rows - 2))))
        w = int(rng.integers(2, max(3, min(16, cols - 2))))
        r, c = rng.integers(0, rows - h + 1), rng.integers(0, cols - w + 1)
        a[r:r + h, c:c + w] = np.maximum(a[r:r + h, c:c + w], rng.uniform(4.0, 25.0))
    if noise:
        a += rng.uniform(0.0, noise, a.shape)
    return a.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize('seed,rows,cols,pixel,threshold,noise', [
    (1, 64, 80, 1.0, 3.0, 0.0),       # boxes on flat ground
    (2, 70, 58, 0.5, 3.0, 0.0),       # 19 px templates (radius 9)
    (3, 48, 48, 1.0, 2.0, 3.5),       # rough terrain: many gradient fallbacks
    (4, 10, 12, 1.0, 3.0, 0.0),       # raster too small for templates: gradients only
    (5, 9, 30, 1.0, 3.0, 0.0),        # first-row column quirk (cols > rows, negative DSM)
])
def test_dsm_walls_match_oracle(rt, seed, rows, cols, pixel, threshold, noise):
    from solweig_amd import walls
    a = _city(seed, rows, cols, noise=noise)
    if seed == 5:
        a[0, :] -= 5.0
    print(f'walls seed {seed} {rows}x{cols} pixel {pixel}', flush=True)
    w, asp = walls.calculate_dsm_walls(a, pixel, threshold)
    rw, rasp = ow.dsm_walls(a, pixel, threshold)
    np.testing.assert_array_equal(w, rw)
    exact = np.isclose(rasp, np.round(rasp), rtol=0, atol=0)
    np.testing.assert_array_equal(asp[exact], rasp[exact])
    np.testing.assert_allclose(asp[~exact], rasp[~exact], rtol=1e-13, atol=1e-11)
    assert np.any(rw > 0)
