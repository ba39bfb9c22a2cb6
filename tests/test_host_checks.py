"""Argument checks of the drop-in entry points that run before any device call (CPU)."""
import os

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
    # two sets of per-hour records (pipelined days), Tmrt + UTCI and two output staging slots
    assert dev == 1_000_000 * (tiling.STATIC_BYTES_PER_PIXEL + 2 * 17 * 24 + 192 * 3)
    assert host == 1_000_000 * 384
    dev, host = tiling.batch_bytes(tiling.DEFAULT_BATCH_PIXELS)
    assert dev < 140e9 and host < 50e9          # the default batch fits one 288 GB device with room


def test_rasters_above_the_pixel_limit_are_refused():
    """The marches address the height planes with 32-bit byte offsets: a domain of more than
    SW_MAX_PIXELS pixels is refused before anything is allocated (the header's limit and the
    bindings' agree)."""
    import re
    from types import SimpleNamespace
    from solweig_amd import model, runtime as rt
    from solweig_amd.static import PreparedDomain
    text = open(os.path.join(os.path.dirname(__file__), '..', 'include', 'solweig_hip.h')).read()
    m = re.search(r'#define SW_MAX_PIXELS \(\(1LL << (\d+)\) - 1\)', text)
    assert m and (1 << int(m.group(1))) - 1 == rt.SW_MAX_PIXELS
    big = np.broadcast_to(np.float32(1.0), (40000, 27000))          # 1.08e9 pixels, no memory
    with pytest.raises(ValueError, match='pixel limit'):
        PreparedDomain(SimpleNamespace(dsm=big, trees=big, dem=big))
    with pytest.raises(ValueError, match='pixel limit'):
        model.prepare_shadow(big, big, big, 1.0)
