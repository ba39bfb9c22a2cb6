"""Seeded synthetic domains and forcing for the five benchmark workloads.

The reference ships no rasters (its example case downloads open data), so every workload here
is generated: flat ground, box buildings, Gaussian-blob canopy, and a sinusoidal summer day of
hourly forcing in the 24-column SUEWS/UMEP met layout that ``series.load_meteorology`` reads
(series.py:175-194; columns per series.py:405-414: Ws=9, RH=10, Ta=11, P=12 kPa, radG=14).

Workloads (names follow their sizes):

* ``c0_plain_cpu``     200x200, 12 boxes 8-30 m, heights U(6, 30) m, no trees, 24 h.
* ``c1_plain_gpu``     1000x1000, boxes to 35 % cover (heights log-normal, median 15 m,
                       sigma_ln 0.6, clipped to 5-60 m), canopy blobs to 15 % cover, 24 h with
                       seeded cloud 0-0.6 scaling the clear-sky global radiation.
* ``c2_batched``       2000x2000 from the same generator, 720 h (30 days from 2025-06-06).
* ``c3_adjoint``       c1 morphology with per-pixel albedo U(0.1, 0.4) and emissivity
                       U(0.90, 0.98) maps.
* ``c4_uneven``        64 domains, sides U_int[256, 1536], latitudes U(0, 60) N.

Walls and aspect are derived here, not by the reference's template matcher (walls.py:24-142 is
static preparation outside the per-hour path); they are fixed inputs that the device path and
the oracle both consume unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
import datetime

import numpy as np

from .sun import met_sun

MET_COLUMNS = 24
COL_WIND, COL_RH, COL_TA, COL_P, COL_RADG = 9, 10, 11, 12, 14


@dataclass
class Domain:
    """One raster domain with its forcing. Rasters are float32 [rows, cols]."""
    dsm: np.ndarray
    trees: np.ndarray
    dem: np.ndarray
    walls: np.ndarray
    aspect: np.ndarray
    met: np.ndarray
    latitude: float
    longitude: float
    utc: float = 0.0
    site_altitude: float = 3.0
    scale: float = 1.0
    landcover: np.ndarray | None = None
    alb_grid: np.ndarray | None = None
    emis_grid: np.ndarray | None = None
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.dsm.shape

    @property
    def location(self):
        return {"latitude": float(self.latitude), "longitude": float(self.longitude),
                "altitude": float(self.site_altitude)}


def wall_heights(dsm, threshold=3.0):
    """Wall height on the ground pixel beside a step: max of the 4 neighbours minus the centre,
    kept where it reaches ``threshold`` (the interior rule of geometry.cu:15-34)."""
    a = np.asarray(dsm, dtype=np.float64)
    out = np.zeros_like(a)
    if a.shape[0] < 3 or a.shape[1] < 3:
        return out
    c = a[1:-1, 1:-1]
    nb = np.maximum(np.maximum(a[:-2, 1:-1], a[2:, 1:-1]), np.maximum(a[1:-1, :-2], a[1:-1, 2:]))
    d = nb - c
    out[1:-1, 1:-1] = np.where(d >= threshold, d, 0.0)
    return out


def wall_aspect(dsm, walls, pixel_size=1.0):
    """Wall aspect in degrees clockwise from north from the DSM gradient (the fallback rule of
    walls.py:130-136); zero off the walls."""
    a = np.asarray(dsm, dtype=np.float64)
    fy, fx = np.gradient(a, pixel_size, pixel_size)
    g = -np.arctan2(fx, fy)
    g[g < 0] += 2 * np.pi
    return np.where(walls > 0, g * (180.0 / np.pi), 0.0)


def _finish_morphology(dsm, trees):
    dsm = dsm.astype(np.float32)
    trees = trees.astype(np.float32)
    walls = wall_heights(dsm).astype(np.float32)
    aspect = wall_aspect(dsm, walls).astype(np.float32)
    return dsm, trees, np.zeros_like(dsm), walls, aspect


def boxes_fixed(rng, rows, cols, count=12, side=(8, 30), height=(6.0, 30.0)):
    dsm = np.zeros((rows, cols), dtype=np.float64)
    for _ in range(count):
        h, w = (int(v) for v in rng.integers(side[0], side[1] + 1, size=2))
        r = int(rng.integers(0, max(1, rows - h + 1)))
        c = int(rng.integers(0, max(1, cols - w + 1)))
        z = float(rng.uniform(*height))
        dsm[r:r + h, c:c + w] = np.maximum(dsm[r:r + h, c:c + w], z)
    return dsm


def boxes_to_cover(rng, rows, cols, cover=0.35, side=(8, 40), median=15.0, sigma=0.6,
                   clip=(5.0, 60.0)):
    dsm = np.zeros((rows, cols), dtype=np.float64)
    target = cover * rows * cols
    covered = 0
    while covered < target:
        h, w = (int(v) for v in rng.integers(side[0], side[1] + 1, size=2))
        h, w = min(h, rows), min(w, cols)
        r = int(rng.integers(0, rows - h + 1))
        c = int(rng.integers(0, cols - w + 1))
        z = float(np.clip(rng.lognormal(np.log(median), sigma), *clip))
        window = dsm[r:r + h, c:c + w]
        covered += int(np.count_nonzero(window == 0))
        np.maximum(window, z, out=window)
    return dsm


def canopy_to_cover(rng, buildings_mask, cover=0.15, peak=(3.0, 15.0), sigma=(2.0, 6.0),
                    cluster=(3, 8), spread=12.0, minimum=3.0):
    rows, cols = buildings_mask.shape
    canopy = np.zeros((rows, cols), dtype=np.float64)
    target = cover * rows * cols
    counted = 0
    ground = ~buildings_mask
    if not ground.any():
        return canopy
    while counted < target:
        cr, cc = rng.uniform(0, rows), rng.uniform(0, cols)
        for _ in range(int(rng.integers(cluster[0], cluster[1] + 1))):
            br = cr + rng.normal(0, spread)
            bc = cc + rng.normal(0, spread)
            p = rng.uniform(*peak)
            s = rng.uniform(*sigma)
            rad = int(np.ceil(3 * s))
            r0, r1 = max(0, int(br) - rad), min(rows, int(br) + rad + 1)
            c0, c1 = max(0, int(bc) - rad), min(cols, int(bc) + rad + 1)
            if r0 >= r1 or c0 >= c1:
                continue
            yy, xx = np.mgrid[r0:r1, c0:c1]
            blob = p * np.exp(-((yy - br) ** 2 + (xx - bc) ** 2) / (2 * s * s))
            win = canopy[r0:r1, c0:c1]
            before = np.count_nonzero((win >= minimum) & ground[r0:r1, c0:c1])
            np.maximum(win, blob, out=win)
            win[~ground[r0:r1, c0:c1]] = 0.0
            after = np.count_nonzero((win >= minimum) & ground[r0:r1, c0:c1])
            counted += after - before
    canopy[canopy < minimum] = 0.0
    return canopy


def day_met(start, days, latitude, longitude, utc, rng_for_day, *, cloudy, altitude=3.0):
    """Hourly forcing for ``days`` consecutive days from ``start`` (a date).

    Ta = 17 + 5 sin(2 pi (h - 9)/24) C, RH = 72.5 - 12.5 sin(...) %, wind 2 m/s, P = 101.3 kPa,
    radG = 800 max(0, sin alt_h) / max_h sin alt_h (x (1 - 0.75 c_h^3.4) when ``cloudy``, with
    hourly cloud c_h ~ U(0, 0.6)); alt_h is the half-step-shifted solar altitude.
    """
    rows = []
    for d in range(days):
        date = start + datetime.timedelta(days=d)
        doy = date.timetuple().tm_yday
        met = np.full((24, MET_COLUMNS), -999.0)
        h = np.arange(24, dtype=np.float64)
        met[:, 0], met[:, 1], met[:, 2], met[:, 3] = date.year, doy, h, 0
        wave = np.sin(2 * np.pi * (h - 9) / 24)
        met[:, COL_TA] = 17 + 5 * wave
        met[:, COL_RH] = 72.5 - 12.5 * wave
        met[:, COL_WIND] = 2.0
        met[:, COL_P] = 101.3
        sun = met_sun(met, {"latitude": latitude, "longitude": longitude, "altitude": altitude}, utc)
        s = np.maximum(0.0, np.sin(np.radians(sun["altitude"])))
        radg = 800.0 * s / max(s.max(), 1e-12)
        if cloudy:
            c = rng_for_day(d).uniform(0.0, 0.6, size=24)
            radg = radg * (1 - 0.75 * c ** 3.4)
        met[:, COL_RADG] = radg
        rows.append(met)
    return np.concatenate(rows, axis=0)


DUBLIN = (53.35, -6.26)


def make_c0(seed=0, size=200, landcover=False):
    rng = np.random.default_rng(seed)
    dsm = boxes_fixed(rng, size, size)
    dsm, trees, dem, walls, aspect = _finish_morphology(dsm, np.zeros_like(dsm))
    met = day_met(datetime.date(2025, 6, 21), 1, *DUBLIN, 0.0, None, cloudy=False)
    lc = np.ones_like(dsm) if landcover else None
    return Domain(dsm, trees, dem, walls, aspect, met, DUBLIN[0], DUBLIN[1], 0.0,
                  landcover=lc, name=f"c0_{size}")


def make_morphology(seed, rows, cols):
    rng = np.random.default_rng(seed)
    dsm = boxes_to_cover(rng, rows, cols)
    canopy = canopy_to_cover(rng, dsm > 0)
    return _finish_morphology(dsm, canopy)


def make_c1(seed=1, size=1000, days=1, start=datetime.date(2025, 6, 21), latitude=DUBLIN[0],
            longitude=DUBLIN[1], met_seed=None, name=None):
    dsm, trees, dem, walls, aspect = make_morphology(seed, size, size)
    base = seed * 100 if met_seed is None else met_seed
    met = day_met(start, days, latitude, longitude, 0.0,
                  lambda d: np.random.default_rng(base + d), cloudy=True)
    return Domain(dsm, trees, dem, walls, aspect, met, latitude, longitude, 0.0,
                  name=name or f"c1_{size}")


def make_c2(seed=2, size=2000, days=30):
    return make_c1(seed=seed, size=size, days=days, start=datetime.date(2025, 6, 6),
                   name=f"c2_{size}_{days}d")


def make_c3(seed=3, size=1000):
    d = make_c1(seed=seed, size=size, name=f"c3_{size}")
    rng = np.random.default_rng(seed + 30)
    d.alb_grid = rng.uniform(0.1, 0.4, size=d.shape).astype(np.float32)
    d.emis_grid = rng.uniform(0.90, 0.98, size=d.shape).astype(np.float32)
    return d


def make_c4(seed=4, count=64, side=(256, 1536), max_side=None):
    rng = np.random.default_rng(seed)
    sides = rng.integers(side[0], side[1] + 1, size=count)
    lats = rng.uniform(0.0, 60.0, size=count)
    out = []
    for i in range(count):
        n = int(sides[i]) if max_side is None else min(int(sides[i]), max_side)
        dsm, trees, dem, walls, aspect = make_morphology(4000 + i, n, n)
        met = day_met(datetime.date(2025, 6, 21), 1, float(lats[i]), 0.0, 0.0,
                      lambda d, i=i: np.random.default_rng(5000 + i + d), cloudy=True)
        out.append(Domain(dsm, trees, dem, walls, aspect, met, float(lats[i]), 0.0, 0.0,
                          name=f"c4_{i}_{n}"))
    return out


def small_domain(seed=7, size=48, trees=True, boxes=5, height=(6.0, 20.0), hours=None,
                 landcover=False, maps=False, latitude=DUBLIN[0]):
    """A small domain for fast parity tests (CPU oracle finishes in seconds)."""
    rng = np.random.default_rng(seed)
    dsm = boxes_fixed(rng, size, size, count=boxes, side=(4, max(5, size // 5)), height=height)
    canopy = np.zeros_like(dsm)
    if trees:
        canopy = canopy_to_cover(rng, dsm > 0, cover=0.12, sigma=(1.5, 3.0), spread=4.0)
    dsm, tr, dem, walls, aspect = _finish_morphology(dsm, canopy)
    met = day_met(datetime.date(2025, 6, 21), 1, latitude, DUBLIN[1], 0.0,
                  lambda d: np.random.default_rng(seed * 100 + d), cloudy=True)
    if hours is not None:
        met = met[list(hours)]
    d = Domain(dsm, tr, dem, walls, aspect, met, latitude, DUBLIN[1], 0.0, name=f"small_{size}")
    if landcover:
        d.landcover = np.ones_like(dsm)
    if maps:
        d.alb_grid = rng.uniform(0.1, 0.4, size=d.shape).astype(np.float32)
        d.emis_grid = rng.uniform(0.90, 0.98, size=d.shape).astype(np.float32)
    return d


WORKLOADS = {
    "c0_plain_cpu": make_c0,
    "c1_plain_gpu": make_c1,
    "c2_batched": make_c2,
    "c3_adjoint": make_c3,
}

__all__ = ["Domain", "make_c0", "make_c1", "make_c2", "make_c3", "make_c4", "small_domain",
           "wall_heights", "wall_aspect", "day_met", "WORKLOADS", "replace"]
