"""Wall heights and aspect from a DSM on the device (drop-in for walls.calculate_dsm_walls,
walls.py:24-142 of the reference).

The rotated line and two-sided templates are built on the host with scipy.ndimage.rotate as the
reference builds them (walls.py:33-56); sw_dsm_walls then computes the heights, the template
search for every wall pixel, the side comparison and the gradient fallback in two kernels.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache

import numpy as np

from . import runtime as rt


def filter_size(pixel_size_m):
    """Template side in pixels (walls.py:24-30)."""
    size = int(np.floor((1.0 / pixel_size_m + 1e-10) * 9))
    if size <= 2:
        size = 3
    elif size != 9 and size % 2 == 0:
        size += 1
    return size


@lru_cache(maxsize=8)
def template_tables(size):
    """(line offsets [L, 2], line starts [181], side offsets [S, 2], side starts [180, 3]) as
    int32: per angle 0..179, the cells of the rotated vertical line (linear interpolation,
    rounded) and of the two halves of the rotated horizontal bar (nearest), relative to the
    centre, in row-major order, with the reference's fixes at angles 30, 45, 135 and 150."""
    from scipy.ndimage import rotate
    r = size // 2
    bar = np.zeros((size, size))
    bar[:, r] = 1
    halves = np.zeros((size, size))
    halves[r, :r] = 1
    halves[r, r + 1:] = 2
    lines, sides, ls, ss = [], [], [0], []
    for angle in range(180):
        rl = np.round(rotate(bar, angle, order=1, reshape=False, mode='nearest'))
        rs = np.round(rotate(halves, angle, order=0, reshape=False, mode='nearest'))
        if angle in (30, 150):
            rs[:, size - 1] = 0
        if angle == 45:                       # direction 225
            rl[0, 0] = rl[size - 1, size - 1] = 1
        if angle == 135:                      # direction 135
            rl[0, size - 1] = rl[size - 1, 0] = 1
        line = np.argwhere(rl != 0) - r
        s1, s2 = np.argwhere(rs == 1) - r, np.argwhere(rs == 2) - r
        base = sum(len(x) for x in sides)
        ss.append((base, base + len(s1), base + len(s1) + len(s2)))
        lines.append(line)
        sides += [s1, s2]
        ls.append(ls[-1] + len(line))
    line = np.concatenate(lines).astype(np.int32)
    side = np.concatenate([s for s in sides if len(s)] or [np.zeros((0, 2))]).astype(np.int32)
    for t in (line, side):
        if len(t) and np.abs(t).max() > r:
            raise AssertionError('template offsets leave the template square')
    return (np.ascontiguousarray(line), np.asarray(ls, dtype=np.int32), np.ascontiguousarray(side),
            np.asarray(ss, dtype=np.int32))


def dsm_walls_device(dsm, pixel_size_m, minimum_wall_height_m=3.0, stream=None):
    """Device FP64 (walls, aspect) arrays of an FP64 DSM (host array or device array)."""
    if isinstance(dsm, rt.DeviceArray):
        d, shape = dsm, dsm.shape
        if d.dtype != np.float64:
            raise ValueError('dsm_walls_device: the DSM must be FP64')
    else:
        a = np.ascontiguousarray(np.asarray(dsm.get() if hasattr(dsm, 'get') else dsm, dtype=np.float64))
        d, shape = rt.DeviceArray.from_host(a), a.shape
    rows, cols = (int(x) for x in shape)
    pixel_size = float(pixel_size_m)
    size = filter_size(pixel_size)
    radius = size // 2
    n_angles = 180 if (rows > 2 * radius + 2 and cols > 2 * radius + 2) else 0
    line, ls, side, ss = template_tables(size)
    up = rt.DeviceArray.from_host
    tabs = [up(x) for x in (line, ls, side, ss)] if n_angles else [None] * 4
    occ = rt.DeviceArray((rows, cols), np.uint8)
    walls = rt.DeviceArray((rows, cols), np.float64)
    aspect = rt.DeviceArray((rows, cols), np.float64)
    step = 1.0 / (1.0 / pixel_size)
    ptr = lambda t: t.ptr if t is not None else None
    f = rt.lib().sw_dsm_walls
    vp, ci = ctypes.c_void_p, ctypes.c_int
    f.argtypes = [vp, ci, ci, ctypes.c_double, ctypes.c_double, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp]
    f.restype = ci
    s = rt._s(stream)
    rt.check(f(d.ptr, rows, cols, float(minimum_wall_height_m), step, radius, n_angles, *(ptr(t) for t in tabs),
               occ.ptr, walls.ptr, aspect.ptr, s))
    rt.check(rt.lib().sw_stream_sync(s))
    del tabs, occ
    return walls, aspect


def calculate_dsm_walls(dsm, pixel_size_m, minimum_wall_height_m=3.0):
    """walls.py:134-142: host FP64 (walls, aspect) of a DSM, computed on the device."""
    w, a = dsm_walls_device(dsm, pixel_size_m, minimum_wall_height_m)
    return w.to_host(), a.to_host()
