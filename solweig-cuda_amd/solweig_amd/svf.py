"""Sky view factors on the device (drop-in for svf.svf_calculator / compute_diffsh, svf.py:81-131).

A sky patch is seen from a pixel exactly as the sun is: the visibility march of a patch is the
shadow march of a sun standing at the patch centre, with the step index starting at 0 and the
first failing step included (svf.py:19-78 builds its tables with the same step rule as the
shadow loop of umep_solweig.py:2286-2358). The per-patch ray tables are therefore the hourly sun
rays of ``sky.shadow_ray`` evaluated at the 153 patch centres. The device marches every
(pixel, patch) pair and keeps the three visibility masks as bits (3 x 153 bits per pixel instead
of three FP32 [rows, cols, 153] arrays), then reduces them to the 15 SVF planes.
"""
from __future__ import annotations

import numpy as np

from . import runtime as rt
from .sky import BAND_COUNTS, patches as patch_centres, shadow_ray

# kernel plane order -> UMEP tuple order (svf.py:113-114)
UMEP_ORDER = (0, 10, 1, 11, 6, 4, 14, 9, 2, 12, 7, 5, 3, 13, 8)
UMEP_NAMES = ('svf', 'svfaveg', 'svfE', 'svfEaveg', 'svfEveg', 'svfN', 'svfNaveg', 'svfNveg',
              'svfS', 'svfSaveg', 'svfSveg', 'svfveg', 'svfW', 'svfWaveg', 'svfWveg')
ANNULI_PER_BAND = 12          # 1-degree annuli summed into each band's weight; the zenith band has 6


def annulus_weight(k):
    """Cosine-weighted solid-angle share of the 1-degree annulus k (1..90 from the zenith), FP64
    in the operation order of svf.py:40-46: sin(1 deg) sin((2 (91 - k) - 1) deg) / (2 pi)."""
    ring = 91. - float(k)
    return (1. / (2. * np.pi)) * np.sin(np.pi / 180.) * np.sin((np.pi * (2. * ring - 1.)) / 180.)


def band_weights(counts=BAND_COUNTS):
    """[bands, 2] FP64: per band the weight of one patch over the full circle (column 0) and over
    a half circle (column 1, the E/S/W/N sector planes), each the running sum over the band's
    annuli in zenith-to-horizon order. Bands span 12 annuli; the last one the remaining 6."""
    out = np.zeros((len(counts), 2))
    first = 1
    for b, n in enumerate(counts):
        last = min(90, first + ANNULI_PER_BAND - 1)
        full = (360. / n) * (np.pi / 180.)
        half = (360. / np.ceil(n / 2.)) * (np.pi / 180.)
        for k in range(first, last + 1):
            w = annulus_weight(k)
            out[b, 0] += full * w
            out[b, 1] += half * w
        first = last + 1
    return out


def sector_bits(azimuth):
    """uint8 per patch: bit d (0 E, 1 S, 2 W, 3 N) is set when the azimuth lies in the half circle
    centred on that compass direction, i.e. in quadrant d or d + 1 of [0, 360) (quadrant q covers
    [90 q, 90 (q + 1)) degrees; svf.py:36-39)."""
    az = np.asarray(azimuth, dtype=np.float64)
    quadrant = (az >= 90.).astype(np.int32) + (az >= 180.) + (az >= 270.)
    bits = np.zeros(az.shape, np.uint8)
    for d in range(4):
        bits |= (((quadrant == d) | (quadrant == (d + 1) % 4)).astype(np.uint8) << d)
    return bits


def ray_tables(rows, cols, amaxvalue, scale):
    """Per-patch ray offsets/heights/lengths, sector bits and band weights of the 153-patch sky
    (the tables of svf.py:19-78): patch p's ray is the sun ray of sky.shadow_ray at the patch
    centre (altitude, azimuth), indices from 0, cut after the first step whose height exceeds
    ``amaxvalue`` or whose shift leaves the raster. Rays are left-aligned in a [153, capacity]
    table, capacity = the longest ray."""
    altitude, azimuth = patch_centres()
    azimuth = np.where(azimuth > 360., azimuth - 360., azimuth)
    rays = [shadow_ray(z, a, scale, amaxvalue, (rows, cols), first=0)[:2] for a, z in zip(altitude, azimuth)]
    lengths = np.array([len(dz) for _, dz in rays], dtype=np.int32)
    capacity = max(1, int(lengths.max()))
    offsets = np.zeros((len(rays), capacity, 2), dtype=np.int32)
    heights = np.zeros((len(rays), capacity), dtype=np.float64)
    for p, (off, dz) in enumerate(rays):
        offsets[p, :len(dz)] = off
        heights[p, :len(dz)] = dz
    counts = np.asarray(BAND_COUNTS, dtype=np.int32)
    return dict(band_offsets=np.concatenate([[0], np.cumsum(counts)]).astype(np.int32),
                sectors=sector_bits(azimuth), weights=band_weights(counts), offsets=offsets,
                heights=heights, lengths=lengths, patches=len(rays), bands=len(counts), capacity=capacity)


class SkyView:
    """Device result of the static sky-view pass: visibility bits and the FP64 SVF planes."""

    def __init__(self, bits, planes, shape):
        self.bits = bits          # DeviceArray uint32 [3, 5, P]
        self.planes = planes      # DeviceArray float64 [16, P] (kernel order + svftotal)
        self.shape = shape

    def umep_tuple(self):
        """Host planes in the order svf_calculator returns them (svf.py:113-114)."""
        if self.planes is None:
            raise ValueError('the FP64 sky-view planes were released after the static preparation '
                             '(PreparedDomain(keep_host=True) keeps a host copy)')
        v = self.planes.to_host().reshape(16, *self.shape)
        return tuple(v[i] for i in UMEP_ORDER)

    def masks(self):
        """Host (shmat, vegmat, vbmat) as float32 [rows, cols, 153] (for parity checks)."""
        return unpack_visibility(self.bits.to_host(), self.shape)


def unpack_visibility(bits, shape, patches=153):
    """uint32 [3, 5, P] bit words -> three float32 [rows, cols, patches] 0/1 masks."""
    b = np.asarray(bits, dtype=np.uint32).reshape(3, rt.SW_VIS_WORDS, -1)
    out = []
    for m in range(3):
        words = np.ascontiguousarray(b[m].T, dtype='<u4')
        u8 = words.view(np.uint8).reshape(words.shape[0], rt.SW_VIS_WORDS * 4)
        bits_m = np.unpackbits(u8, axis=1, bitorder='little')[:, :patches]
        out.append(bits_m.reshape(*shape, patches).astype(np.float32))
    return tuple(out)


def sky_view(dsm_dev, trees_dev, dsm_raise, shape, amaxvalue, scale, transmissivity=0.03, stream=None,
             events=None, hmax=None, march_trees=None):
    """Run the static visibility march and band reduction for one domain. ``events``: three
    runtime.Event recorded before the march, between the passes and after the reduction.
    ``hmax`` (device array of static.height_bound, heights dsm + raise >= 0) selects the windowed
    sun-ray march; ``march_trees`` is its optional canopy plane (static.march_trees)."""
    rows, cols = shape
    t = ray_tables(rows, cols, float(amaxvalue), float(scale))
    P = rows * cols
    offsets, heights, capacity = t['offsets'], t['heights'], int(t['capacity'])
    if hmax is not None:
        # the march loads 8 steps at a time: rays start at multiples of 8 steps
        capacity = -(-capacity // 8) * 8
        offsets = np.zeros((t['patches'], capacity, 2), np.int32)
        heights = np.zeros((t['patches'], capacity), np.float64)
        offsets[:, :t['capacity']] = t['offsets']
        heights[:, :t['capacity']] = t['heights']
    off = rt.DeviceArray.from_host(offsets, np.int32)
    hts = rt.DeviceArray.from_host(heights, np.float64)
    lens = rt.DeviceArray.from_host(t['lengths'], np.int32)
    sec = rt.DeviceArray.from_host(t['sectors'], np.uint8)
    bo = rt.DeviceArray.from_host(t['band_offsets'], np.int32)
    w = rt.DeviceArray.from_host(t['weights'], np.float64)
    bits = rt.DeviceArray((3, rt.SW_VIS_WORDS, P), np.uint32)
    planes = rt.DeviceArray((16, P), np.float64)
    s = rt._s(stream)
    if events:
        events[0].record(stream)
    rt.call('sw_svf_visibility', dsm_dev.ptr, trees_dev.ptr, 0 if march_trees is None else march_trees.ptr,
            0 if hmax is None else hmax.ptr, -(-cols // rt.HMAX_CELL), float(dsm_raise), rows, cols, off.ptr,
            hts.ptr, lens.ptr, capacity, int(t['patches']), bits.ptr, s)
    if events:
        events[1].record(stream)
    rt.call('sw_svf_reduce', bits.ptr, dsm_dev.ptr, trees_dev.ptr, float(dsm_raise), rows, cols, sec.ptr,
            bo.ptr, w.ptr, int(t['bands']), int(t['patches']), float(transmissivity), planes.ptr, s)
    if events:
        events[2].record(stream)
    rt.call('sw_stream_sync', s)
    return SkyView(bits, planes, shape)


def svf_calculator(patch_option, amaxvalue, a, vegdem, vegdem2, bush, scale,
                   angle_batch=None, transmissivity=0.03):
    """Drop-in of svf.svf_calculator (svf.py:81-114) for host inputs: returns the 15 UMEP-ordered
    FP64 planes followed by (vegmat, vbmat, shmat, svftotal) as host arrays.

    ``a`` is the DSM; ``vegdem`` the canopy DSM (trees + a, zero where no trees). The device
    rebuilds both vegetation DSMs from FP32 rasters, so a and vegdem - a must be FP32-exact."""
    if int(patch_option) != 2:
        raise ValueError('only patch_option 2 (153 patches) is supported')
    a = np.asarray(a, dtype=np.float64)
    top = np.asarray(vegdem, dtype=np.float64)
    trees = np.where(top != 0, top - a, 0.0)
    dsm32, trees32 = a.astype(np.float32), trees.astype(np.float32)
    if not (np.array_equal(dsm32.astype(np.float64), a) and
            np.array_equal(np.where(top != 0, trees32.astype(np.float64) + a, 0.0), top)):
        raise ValueError('svf_calculator: DSM and canopy heights must be FP32-exact rasters')
    d = rt.DeviceArray.from_host(dsm32)
    tr = rt.DeviceArray.from_host(trees32)
    hmax = mt = None
    if a.min() >= 0.0:
        # heights >= 0: the windowed sun-ray march (a DSM below 0 keeps the plain per-pixel march,
        # whose off-raster rule depends on the sign of the height)
        from .static import height_bound, march_trees
        hmax = rt.DeviceArray.from_host(height_bound(a, trees32))
        m = march_trees(a, trees32)
        mt = None if m is None else rt.DeviceArray.from_host(m)
    sv = sky_view(d, tr, 0.0, a.shape, amaxvalue, scale, transmissivity, hmax=hmax, march_trees=mt)
    planes = sv.planes.to_host().reshape(16, *a.shape)
    shm, vgm, vbm = sv.masks()
    return tuple(planes[i] for i in UMEP_ORDER) + (vgm, vbm, shm, planes[15])


def compute_diffsh(shmat, vegshmat, transmissivity):
    """diffsh = shmat - (1 - vegshmat)(1 - t) in FP32 (svf.py:117-131); host arrays."""
    one = np.float32(1.0)
    return (np.asarray(shmat, np.float32) - (one - np.asarray(vegshmat, np.float32))
            * (one - np.float32(transmissivity))).astype(np.float32)
