"""153-patch sky view factors and visibility masks (oracle; test infrastructure only).

Restates the reference's SVF path: ray tables (svf.py:19-78), per (pixel, patch) visibility
march (geometry.cu:66-119), band-weighted reduction (geometry.cu:121-159) and the reorder to
the UMEP tuple (svf.py:113-114). Comparisons are the exact FP64 ones of geometry.cu:74.
"""
from __future__ import annotations

import numpy as np

COUNTS = (31, 30, 28, 24, 19, 13, 7, 1)
EDGES = (0, 12, 24, 36, 48, 60, 72, 84, 90)
ALTITUDES = (6, 18, 30, 42, 54, 66, 78, 90)
STARTS = (0, 4, 2, 5, 8, 0, 10, 0)
UMEP_ORDER = (0, 10, 1, 11, 6, 4, 14, 9, 2, 12, 7, 5, 3, 13, 8)


def ray_tables(rows, cols, amaxvalue, scale):
    """Per-patch offsets/heights/lengths, sector bits and band weights (svf.py:19-78)."""
    counts = np.asarray(COUNTS)
    band = np.repeat(np.arange(len(counts)), counts)
    band_offsets = np.r_[0, np.cumsum(counts)]
    azimuth = np.concatenate([np.arange(n) * (360. / n) + s for n, s in zip(counts, STARTS)])
    azimuth[azimuth > 360.] -= 360.
    altitude = np.asarray(ALTITUDES)[band].astype(np.float64)
    sectors = ((azimuth >= 0) & (azimuth < 180)).astype(np.uint8)
    sectors |= ((azimuth >= 90) & (azimuth < 270)).astype(np.uint8) << 1
    sectors |= ((azimuth >= 180) & (azimuth < 360)).astype(np.uint8) << 2
    sectors |= ((azimuth >= 270) | (azimuth < 90)).astype(np.uint8) << 3
    weights = np.zeros((len(counts), 2))
    for b, n in enumerate(counts):
        for k in np.arange(EDGES[b] + 1, EDGES[b + 1] + 1):
            ann = 91. - k
            w = (1. / (2. * np.pi)) * np.sin(np.pi / 180.) * np.sin((np.pi * (2. * ann - 1.)) / 180.)
            weights[b, 0] += ((360. / n) * (np.pi / 180.)) * w
            weights[b, 1] += ((360. / np.ceil(n / 2.)) * (np.pi / 180.)) * w
    cap = max(rows, cols) + 1
    idx = np.arange(cap, dtype=np.float64)
    P = len(azimuth)
    offsets = np.zeros((P, cap, 2), dtype=np.int64)
    heights = np.zeros((P, cap))
    lengths = np.zeros(P, dtype=np.int64)
    for p in range(P):
        if amaxvalue < 0:
            continue
        az, el = azimuth[p] * (np.pi / 180.), altitude[p] * (np.pi / 180.)
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
        n = int(bad[0]) + 1 if len(bad) else cap
        lengths[p] = n
        offsets[p, :n, 0], offsets[p, :n, 1] = dx[:n], dy[:n]
        heights[p, :n] = dz[:n]
    return dict(azimuth=azimuth, altitude=altitude, sectors=sectors, weights=weights,
                band_offsets=band_offsets, offsets=offsets, heights=heights, lengths=lengths)


def _shifted(src, dx, dy, fill=0.0):
    """out[r, c] = src[r + dx, c + dy] where inside, ``fill`` elsewhere; also the inside mask."""
    rows, cols = src.shape
    out = np.full(src.shape, fill, dtype=src.dtype)
    inside = np.zeros(src.shape, dtype=bool)
    r0, r1 = max(0, -dx), min(rows, rows - dx)
    c0, c1 = max(0, -dy), min(cols, cols - dy)
    if r0 < r1 and c0 < c1:
        out[r0:r1, c0:c1] = src[r0 + dx:r1 + dx, c0 + dy:c1 + dy]
        inside[r0:r1, c0:c1] = True
    return out, inside


def visibility(a, top, bottom, bush, tables):
    """Per-(pixel, patch) masks (geometry.cu:77-119). Returns shmat, vegmat, vbmat as float32
    [rows, cols, patches] with 1 = sky visible / not vegetation-blocked."""
    rows, cols = a.shape
    P = len(tables['lengths'])
    shm = np.empty((rows, cols, P), dtype=np.float32)
    vgm = np.empty_like(shm)
    vbm = np.empty_like(shm)
    for p in range(P):
        sh = np.zeros(a.shape, dtype=bool)
        veg = bush > 1
        ever = np.zeros(a.shape, dtype=bool)
        done = np.zeros(a.shape, dtype=bool)
        for n in range(int(tables['lengths'][p])):
            if done.all():
                break
            dx, dy = (int(v) for v in tables['offsets'][p, n])
            dz = tables['heights'][p, n]
            dzp = tables['heights'][p, max(n - 1, 0)]
            A, inside = _shifted(a, dx, dy)
            Tt, _ = _shifted(top, dx, dy)
            Bt, _ = _shifted(bottom, dx, dy)
            live = ~done
            out = live & ~inside
            if out.any():
                s_out = a < 0.0
                sh = np.where(out, s_out, sh)
                veg = np.where(out & s_out, False, veg)
                ever = np.where(out, ever | veg, ever)
                done |= out
            lin = live & inside
            hit = (A - dz) > a
            cnt = ((Tt - dz) > a).astype(int) + ((Bt - dz) > a) + ((Tt - dzp) > a) + ((Bt - dzp) > a)
            newveg = veg | ((cnt > 0) & (cnt < 4))
            sh = np.where(lin, hit, sh)
            veg = np.where(lin, np.where(hit, False, newveg), veg)
            ever = np.where(lin & ~hit, ever | newveg, ever)
            done |= lin & hit
        shm[:, :, p] = 1.0 - sh.astype(np.float32)
        vgm[:, :, p] = 1.0 - veg.astype(np.float32)
        vbm[:, :, p] = 1.0 - (ever.astype(np.float32) - veg.astype(np.float32))
    return shm, vgm, vbm


def reduce(shm, vgm, vbm, bottom, tables, trans=0.03):
    """Band-weighted SVF planes (geometry.cu:121-159): 15 planes + svftotal, FP64."""
    rows, cols, P = shm.shape
    v = np.zeros((15, rows, cols))
    bo = tables['band_offsets']
    sec = tables['sectors']
    for b in range(len(bo) - 1):
        cnt = np.zeros((15, rows, cols), dtype=np.int64)
        for j in range(bo[b], bo[b + 1]):
            s = shm[:, :, j].astype(np.int64)
            g = vgm[:, :, j].astype(np.int64)
            u = vbm[:, :, j].astype(np.int64)
            cnt[0] += s
            cnt[5] += g
            cnt[10] += u
            for d in range(4):
                if sec[j] & (1 << d):
                    cnt[1 + d] += s
                    cnt[6 + d] += g
                    cnt[11 + d] += u
        wf, wa = tables['weights'][b]
        for grp in range(3):
            v[grp * 5] += wf * cnt[grp * 5]
            for d in range(1, 5):
                v[grp * 5 + d] += wa * cnt[grp * 5 + d]
    eps = 3.0459e-4
    v[2] += eps
    v[3] += eps
    last = np.where(bottom == 0.0, eps, 0.0)
    for i in (7, 8, 12, 13):
        v[i] += last
    v = np.minimum(v, 1.0)
    total = v[0] - (1.0 - v[5]) * (1.0 - trans)
    return v, total


def svf_calculator(amaxvalue, a, vegdem, vegdem2, bush, scale, trans=0.03):
    """The reference's svf_calculator (svf.py:81-114): returns the UMEP 15-tuple order followed by
    (vegmat, vbmat, shmat, svftotal)."""
    rows, cols = a.shape
    t = ray_tables(rows, cols, float(amaxvalue), float(scale))
    shm, vgm, vbm = visibility(a, vegdem, vegdem2, bush, t)
    v, total = reduce(shm, vgm, vbm, vegdem2, t, trans)
    return tuple(v[i] for i in UMEP_ORDER) + (vgm, vbm, shm, total)


def diffsh(shmat, vegmat, trans=0.03):
    """diffsh = sh - (1 - veg)(1 - trans) in FP32 (geometry.cu:161-172)."""
    return (shmat - (np.float32(1.0) - vegmat) * (np.float32(1.0) - np.float32(trans))).astype(np.float32)
