"""Argument checks of the drop-in entry points that run before any device call (CPU)."""
import numpy as np
import pytest

from solweig_amd import batched, model, synthetic, tiling


def test_prepare_shadow_refuses_heights_below_zero():
    a = np.zeros((8, 8))
    a[2, 3] = -1.0
    z = np.zeros_like(a)
    with pytest.raises(ValueError, match='below 0'):
        model.prepare_shadow(a, z, z, 1.0)


def test_engine_refuses_mixed_pixel_sizes():
    d1 = synthetic.small_domain(seed=1, size=12)
    d2 = synthetic.small_domain(seed=2, size=14)
    d2.scale = 0.5
    with pytest.raises(ValueError, match='one pixel size'):
        batched.Engine([d1, d2])


def test_batch_bytes_model():
    dev, host = tiling.batch_bytes(1_000_000, n_hours=24, slots=2)
    assert dev == 1_000_000 * (tiling.STATIC_BYTES_PER_PIXEL + 9 * 24 + 192 * 3)
    assert host == 1_000_000 * 384
    dev, host = tiling.batch_bytes(tiling.DEFAULT_BATCH_PIXELS)
    assert dev < 120e9 and host < 50e9          # the default batch fits one 288 GB device with room
