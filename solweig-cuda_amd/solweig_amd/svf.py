"""Sky view factors on the device (drop-in for svf.svf_calculator / compute_diffsh, svf.py:81-131).

Ray tables are built on the host exactly as svf.py:19-78 builds them (round-half-even shifts,
FP64 heights, the first failing step included); the device marches every (pixel, patch) pair
and keeps the three visibility masks as bits (3 x 153 bits per pixel instead of three FP32
[rows, cols, 153] arrays), then reduces them to the 15 SVF planes.
"""
from __future__ import annotations

import numpy as np

from . import runtime as rt
from .sky import BAND_ALTITUDES, BAND_COUNTS, BAND_STARTS

_EDGES = (0, 12, 24, 36, 48, 60, 72, 84, 90)
# kernel plane order -> UMEP tuple order (svf.py:113-114)
UMEP_ORDER = (0, 10, 1, 11, 6, 4, 14, 9, 2, 12, 7, 5, 3, 13, 8)
UMEP_NAMES = ('svf', 'svfaveg', 'svfE', 'svfEaveg', 'svfEveg', 'svfN', 'svfNaveg', 'svfNveg',
              'svfS', 'svfSaveg', 'svfSveg', 'svfveg', 'svfW', 'svfWaveg', 'svfWveg')


def ray_tables(rows, cols, amaxvalue, scale):
    """Per-patch ray offsets/heights/lengths, sector bits and band weights (svf.py:19-78)."""
    counts = np.asarray(BAND_COUNTS, dtype=np.int32)
    band = np.repeat(np.arange(len(counts), dtype=np.int32), counts)
    band_offsets = np.r_[0, np.cumsum(counts)].astype(np.int32)
    azimuth = np.concatenate([np.arange(n) * (360. / n) + s for n, s in zip(counts, BAND_STARTS)])
    azimuth[azimuth > 360.] -= 360.
    altitude = np.asarray(BAND_ALTITUDES)[band].astype(np.float64)
    sectors = ((azimuth >= 0) & (azimuth < 180)).astype(np.uint8)
    sectors |= (((azimuth >= 90) & (azimuth < 270)).astype(np.uint8) << 1)
    sectors |= (((azimuth >= 180) & (azimuth < 360)).astype(np.uint8) << 2)
    sectors |= (((azimuth >= 270) | (azimuth < 90)).astype(np.uint8) << 3)
    weights = np.zeros((len(counts), 2), dtype=np.float64)
    for b, n in enumerate(counts):
        for k in np.arange(_EDGES[b] + 1, _EDGES[b + 1] + 1):
            annulus = 91. - k
            w = (1. / (2. * np.pi)) * np.sin(np.pi / 180.) * np.sin((np.pi * (2. * annulus - 1.)) / 180.)
            weights[b, 0] += ((360. / n) * (np.pi / 180.)) * w
            weights[b, 1] += ((360. / np.ceil(n / 2.)) * (np.pi / 180.)) * w
    patches = len(azimuth)
    capacity = max(rows, cols) + 1
    idx = np.arange(capacity, dtype=np.float64)
    offsets = np.zeros((patches, capacity, 2), dtype=np.int32)
    heights = np.zeros((patches, capacity), dtype=np.float64)
    lengths = np.zeros(patches, dtype=np.int32)
    for p, (azi, alt) in enumerate(zip(azimuth, altitude)):
        if amaxvalue < 0:
            continue
        az, el = azi * (np.pi / 180.), alt * (np.pi / 180.)
        sn, cs, tn = np.sin(az), np.cos(az), np.tan(az)
        if np.pi / 4 <= az < 3 * np.pi / 4 or 5 * np.pi / 4 <= az < 7 * np.pi / 4:
            dy = np.sign(sn) * idx
            dx = -np.sign(cs) * np.abs(np.round(idx / tn))
            ds = np.abs(1. / sn)
        else:
            dy = np.sign(sn) * np.abs(np.round(idx * tn))
            dx = -np.sign(cs) * idx
            ds = np.abs(1. / cs)
        dz = (ds * idx) * (np.tan(el) / scale)
        bad = np.flatnonzero((dz > amaxvalue) | (np.abs(dx) >= rows) | (np.abs(dy) >= cols))
        n = int(bad[0]) + 1 if len(bad) else capacity
        lengths[p] = n
        offsets[p, :n, 0], offsets[p, :n, 1] = dx[:n], dy[:n]
        heights[p, :n] = dz[:n]
    capacity = max(1, int(lengths.max()))
    return dict(band_offsets=band_offsets, sectors=sectors, weights=weights,
                offsets=np.ascontiguousarray(offsets[:, :capacity]),
                heights=np.ascontiguousarray(heights[:, :capacity]), lengths=lengths,
                patches=patches, bands=len(counts), capacity=capacity)


class SkyView:
    """Device result of the static sky-view pass: visibility bits and the FP64 SVF planes."""

    def __init__(self, bits, planes, shape):
        self.bits = bits          # DeviceArray uint32 [3, 5, P]
        self.planes = planes      # DeviceArray float64 [16, P] (kernel order + svftotal)
        self.shape = shape

    def umep_tuple(self):
        """Host planes in the order svf_calculator returns them (svf.py:113-114)."""
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
             events=None):
    """Run the static visibility march and band reduction for one domain. ``events``: three
    runtime.Event recorded before the march, between the passes and after the reduction."""
    rows, cols = shape
    t = ray_tables(rows, cols, float(amaxvalue), float(scale))
    P = rows * cols
    off = rt.DeviceArray.from_host(t['offsets'], np.int32)
    hts = rt.DeviceArray.from_host(t['heights'], np.float64)
    lens = rt.DeviceArray.from_host(t['lengths'], np.int32)
    sec = rt.DeviceArray.from_host(t['sectors'], np.uint8)
    bo = rt.DeviceArray.from_host(t['band_offsets'], np.int32)
    w = rt.DeviceArray.from_host(t['weights'], np.float64)
    bits = rt.DeviceArray((3, rt.SW_VIS_WORDS, P), np.uint32)
    planes = rt.DeviceArray((16, P), np.float64)
    s = rt._s(stream)
    if events:
        events[0].record(stream)
    rt.call('sw_svf_visibility', dsm_dev.ptr, trees_dev.ptr, float(dsm_raise), rows, cols, off.ptr, hts.ptr,
            lens.ptr, int(t['capacity']), int(t['patches']), bits.ptr, s)
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
    sv = sky_view(d, tr, 0.0, a.shape, amaxvalue, scale, transmissivity)
    planes = sv.planes.to_host().reshape(16, *a.shape)
    shm, vgm, vbm = sv.masks()
    return tuple(planes[i] for i in UMEP_ORDER) + (vgm, vbm, shm, planes[15])


def compute_diffsh(shmat, vegshmat, transmissivity):
    """diffsh = shmat - (1 - vegshmat)(1 - t) in FP32 (svf.py:117-131); host arrays."""
    one = np.float32(1.0)
    return (np.asarray(shmat, np.float32) - (one - np.asarray(vegshmat, np.float32))
            * (one - np.float32(transmissivity))).astype(np.float32)
