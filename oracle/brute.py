"""Per-pixel loop form of the sun-ray shadow march (oracle; test infrastructure only).

Independent of the shifted-array form in ``umep.shadows``: each pixel walks its own ray with the
step rule of umep_solweig.py:2286-2358 (out-of-domain samples read 0, the loop runs until the
first step whose dz exceeds amaxvalue or whose shift leaves the raster, inclusive). Used on
tiny rasters to pin the vectorised form.
"""
from __future__ import annotations

import numpy as np

from .umep import march_offsets


def ray(azimuth_deg, altitude_deg, scale, amaxvalue, shape):
    az = azimuth_deg * (np.pi / 180.)
    al = altitude_deg * (np.pi / 180.)
    s, c, t = np.sin(az), np.cos(az), np.tan(az)
    tanalt = np.tan(al) / scale
    steps = []
    dx = dy = dz = 0
    index = 0
    while (amaxvalue >= dz) and (np.abs(dx) < shape[0]) and (np.abs(dy) < shape[1]):
        dx, dy, ydom = march_offsets(az, index, t, np.sign(s), np.sign(c))
        ds = np.abs(1 / s) if ydom else np.abs(1 / c)
        dz = (ds * index) * tanalt
        steps.append((int(dx), int(dy), dz))
        index += 1
    return steps


def shadow_masks(a, vegdem, vegdem2, bush, azimuth_deg, altitude_deg, scale, amaxvalue):
    """(vegsh, sh, vbshvegsh) with 1 = sunlit, by an explicit per-pixel loop."""
    rows, cols = a.shape
    steps = ray(azimuth_deg, altitude_deg, scale, amaxvalue, a.shape)
    vegsh = np.ones(a.shape)
    sh = np.ones(a.shape)
    vb = np.ones(a.shape)
    for r in range(rows):
        for c in range(cols):
            z = a[r, c]
            f = z
            veg = bush[r, c] > 1
            ever = False
            prev = 0.0
            for dx, dy, dz in steps:
                rr, cc = r + dx, c + dy
                inside = 0 <= rr < rows and 0 <= cc < cols
                top = vegdem[rr, cc] if inside else 0.0
                bot = vegdem2[rr, cc] if inside else 0.0
                bld = a[rr, cc] if inside else 0.0
                f = max(f, (bld - dz) if inside else 0.0)
                blocked = f > z
                n = int((top - dz if inside else 0.0) > z) + int((bot - dz if inside else 0.0) > z) \
                    + int((top - prev if inside else 0.0) > z) + int((bot - prev if inside else 0.0) > z)
                prev = dz
                veg = veg or (0 < n < 4)
                if blocked:
                    veg = False
                ever = ever or veg
            sh[r, c] = 0.0 if f > z else 1.0
            vegsh[r, c] = 0.0 if veg else 1.0
            vb[r, c] = 1.0 - (float(ever) - float(veg))
    return vegsh, sh, vb


def sky_masks(a, top, bottom, bush, scale, amaxvalue):
    """(shmat, vegmat, vbmat) float32 [rows, cols, 153], 1 = sky visible / not seen through
    canopy, by an explicit loop over pixels and sky patches (geometry.cu:77-119).

    Each patch centre of create_patches (umep_solweig.py:1737-1786) is treated as a sun at that
    azimuth and altitude: the pixel walks ``ray`` (index from 0, first failing step included) and

    * stops at the first step whose building sample minus the step height is above the pixel
      (sky hidden, canopy flag cleared);
    * stops at the first step outside the raster, where the sky is hidden only if the pixel lies
      below 0 (the zero fill of the shifted arrays);
    * otherwise sets the canopy flag when 1 to 3 of the 4 tests {canopy top, trunk top} minus
      {this step's, the previous step's height} are above the pixel; it starts as bush > 1.

    ``ever`` records that the canopy flag was set at some step; vbmat = 1 - (ever - veg)."""
    from .umep import create_patches
    rows, cols = a.shape
    alt, azi = create_patches(2)
    azi = np.where(azi > 360., azi - 360., azi)
    n_patch = len(alt)
    shm = np.ones((rows, cols, n_patch), np.float32)
    vgm = np.ones((rows, cols, n_patch), np.float32)
    vbm = np.ones((rows, cols, n_patch), np.float32)
    for k in range(n_patch):
        steps = ray(azi[k], alt[k], scale, amaxvalue, a.shape)
        for r in range(rows):
            for c in range(cols):
                z = a[r, c]
                sh = False
                veg = bool(bush[r, c] > 1)
                ever = False
                prev = 0.0
                for dx, dy, dz in steps:
                    rr, cc = r + dx, c + dy
                    if not (0 <= rr < rows and 0 <= cc < cols):
                        sh = bool(z < 0.0)
                        veg = veg and not sh
                        ever = ever or veg
                        break
                    if a[rr, cc] - dz > z:
                        sh, veg = True, False
                        break
                    above = [v - h > z for v in (top[rr, cc], bottom[rr, cc]) for h in (dz, prev)]
                    veg = veg or 0 < sum(above) < 4
                    ever = ever or veg
                    prev = dz
                shm[r, c, k] = 0.0 if sh else 1.0
                vgm[r, c, k] = 0.0 if veg else 1.0
                vbm[r, c, k] = 0.0 if (ever and not veg) else 1.0
    return shm, vgm, vbm
