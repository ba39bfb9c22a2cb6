"""Wall heights and wall aspect from a DSM (oracle; test infrastructure only).

Restates calculate_dsm_walls (walls.py:24-142 and geometry.cu:15-63 of the reference):
- height: the largest of the four edge neighbours minus the pixel, kept where >= threshold,
  zero on the raster border except the first-row quirk of geometry.cu:30-33;
- aspect: each wall pixel picks, among 180 rotated line templates (scipy.ndimage.rotate of a
  vertical line, linear interpolation, rounded), the first angle whose template covers the most
  wall pixels; the rotated two-sided template (nearest) then compares the DSM sums on its two
  halves to choose between the two facing directions 270 - angle and 90 - angle (mod 360);
- fallback: wall pixels left at aspect 0 (near the border, or facing exactly north) take the
  direction of the DSM gradient, -arctan2(d/dx, d/dy) in degrees on [0, 360).

Third-party arithmetic: scipy.ndimage.rotate (SciPy 1.15.3 here) builds the templates exactly as
the reference's walls.py:33-56 calls it. Parity of the templates with the reference's SciPy is
unpinned; the tests pin the rest by closed forms (boxes on flat ground).
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np


def filter_size(pixel_size_m):
    """Template side in pixels (walls.py:24-30)."""
    size = int(np.floor((1.0 / pixel_size_m + 1e-10) * 9))
    if size <= 2:
        size = 3
    elif size != 9 and size % 2 == 0:
        size += 1
    return size


@lru_cache(maxsize=8)
def templates(size):
    """Per angle 0..179: (line, side 1, side 2) offsets (row, col) relative to the centre
    (walls.py:33-56)."""
    from scipy.ndimage import rotate
    r = size // 2
    line = np.zeros((size, size))
    line[:, r] = 1
    sides = np.zeros((size, size))
    sides[r, :r] = 1
    sides[r, r + 1:] = 2
    out = []
    for angle in range(180):
        rl = np.round(rotate(line, angle, order=1, reshape=False, mode='nearest'))
        rs = np.round(rotate(sides, angle, order=0, reshape=False, mode='nearest'))
        if angle in (30, 150):
            rs[:, size - 1] = 0
        if 270 - angle == 225:
            rl[0, 0] = rl[size - 1, size - 1] = 1
        if 270 - angle == 135:
            rl[0, size - 1] = rl[size - 1, 0] = 1
        out.append(tuple(np.argwhere(m).astype(np.int64) - r for m in (rl != 0, rs == 1, rs == 2)))
    return tuple(out)


def wall_heights(dsm, threshold):
    """geometry.cu:15-35."""
    a = np.asarray(dsm, dtype=np.float64)
    rows, cols = a.shape
    h = np.zeros_like(a)
    if rows > 2 and cols > 2:
        nb = np.fmax(np.fmax(a[:-2, 1:-1], a[2:, 1:-1]), np.fmax(a[1:-1, :-2], a[1:-1, 2:]))
        d = nb - a[1:-1, 1:-1]
        h[1:-1, 1:-1] = np.where(d >= threshold, d, 0.0)
    if rows > 1:
        c = np.arange(cols)
        sel = (c >= rows) & (c < cols - 1)
        d = -a[0, sel]
        h[0, sel] = np.where(d >= threshold, d, 0.0)
    return h


def wall_aspect(walls, dsm, pixel_size):
    """walls.py:79-131 with geometry.cu:37-63 (the direction search)."""
    surface = np.asarray(dsm, dtype=np.float64)
    mask = walls > 0
    aspect = np.zeros(surface.shape)
    if not np.any(mask):
        return aspect
    rows, cols = surface.shape
    size = filter_size(pixel_size)
    r = size // 2
    if rows > 2 * r + 2 and cols > 2 * r + 2:
        tpl = templates(size)
        occ = mask.astype(np.int64)
        for p in np.flatnonzero(mask):
            i, j = divmod(int(p), cols)
            if i < r or i >= rows - r - 2 or j < r or j >= cols - r - 2:
                continue
            best, angle = 0, -1
            for ang in range(180):
                off = tpl[ang][0]
                score = int(occ[i + off[:, 0], j + off[:, 1]].sum())
                if score > best:
                    best, angle = score, ang
            if angle < 0:
                continue
            _, s1, s2 = tpl[angle]
            first = surface[i + s1[:, 0], j + s1[:, 1]].sum(dtype=np.float64)
            second = surface[i + s2[:, 0], j + s2[:, 1]].sum(dtype=np.float64)
            direction = 270.0 - angle
            if first > second:
                direction -= 180.0
            if direction < 0:
                direction += 360.0
            aspect[i, j] = direction
    step = 1.0 / (1.0 / pixel_size)
    fy, fx = np.gradient(surface, step, step)
    g = -np.arctan2(fx, fy)
    g[g < 0] += 2 * np.pi
    unresolved = mask & (aspect == 0)
    aspect += unresolved * (g / (math.pi / 180.0))
    return aspect


def dsm_walls(dsm, pixel_size_m, minimum_wall_height_m=3.0):
    """calculate_dsm_walls (walls.py:134-142): (walls, aspect), FP64."""
    walls = wall_heights(dsm, float(minimum_wall_height_m))
    return walls, wall_aspect(walls, dsm, float(pixel_size_m))
