"""Static preparation of a domain: host derivations, sky view on the device, device planes.

Follows series._prepare_spatial (series.py:252-349): DSM/DEM raised to >= 0, trees clipped at 0,
ground mask from DSM - DEM < 2 m, land-cover maps (Tgmaps_v1) or the default constants, amaxvalue,
the SVF pass and the derived planes svfbuveg, asvf, svfalfa. The device keeps only what the hour
kernels read: FP32 rasters, the visibility bits, FP64 tan(asvf)/tan(svfalfa) and FP32 maps.
"""
from __future__ import annotations

import numpy as np

from . import runtime as rt
from .sky import CONSTANTS, DEG, _geometry
from .svf import sky_view

LANDCOVER = {  # code: (albedo, emissivity, TgK, Tstart, TmaxLST) of landcoverclasses_2016a.txt
    2: (0.18, 0.95, 0.58, -9.78, 15.0), 1: (0.18, 0.95, 0.58, -9.78, 15.0),
    0: (0.20, 0.95, 0.37, -3.41, 15.0), 7: (0.05, 0.98, 0.00, 0.00, 12.0),
    5: (0.16, 0.94, 0.21, -3.38, 14.0), 6: (0.25, 0.94, 0.33, -3.01, 14.0),
    99: (0.20, 0.90, 0.58, -3.41, 15.0)}


def normalise_landcover(values):
    """series.py:505-513: non-integer or out-of-range codes -> 6, codes 3/4 -> 5."""
    src = np.asarray(values, dtype=float)
    r = np.rint(src)
    bad = (~np.isfinite(src)) | (~np.isclose(src, r)) | (r < 1) | (r > 7)
    r[bad] = 6
    r[(r == 3) | (r == 4)] = 5
    return r.astype(np.float32)


def sky_sum_table():
    """Per patch (w sin, w cos, g w sin, g w cos) of the day patch geometry, w the steradian and
    g = exp(0.308 (1.7 - 1/cos(zenith))) the Martin-Berdahl band factor (umep_solweig.py:526-527)."""
    alt, _, ster, wcos, wsin = _geometry(True)[:5]
    g = np.exp(0.308 * (1.7 - (1 / np.cos((90 - alt) * DEG))))
    return np.ascontiguousarray(np.column_stack((wsin, wcos, g * wsin, g * wcos)), dtype=np.float32)


def sky_sums(vis, P, stream=None):
    """Static per-pixel longwave patch sums [SKY_SUMS, P] from the visibility bits."""
    tab = rt.DeviceArray.from_host(sky_sum_table())
    out = rt.DeviceArray((rt.SKY_SUMS, P), np.float32)
    rt.call('sw_sky_sums', vis.ptr, int(P), tab.ptr, out.ptr, rt._s(stream))
    rt.call('sw_stream_sync', rt._s(stream))
    return out


def march_trees(a, trees):
    """The canopy plane of the sun-ray march (sw_domain.march_trees): trees where the FP64
    canopy top trees + a differs from a, -inf where it does not (series.py:282-287 zeroes those
    tops), or None when some pixel has 0.25 * trees + a == a but trees + a != a (then the march
    forms the tops itself). ``a`` is the FP64 DSM raised to >= 0."""
    a = np.asarray(a, dtype=np.float64)
    t32 = np.asarray(trees, dtype=np.float32)
    t = t32.astype(np.float64)
    no_top = (t + a) == a
    no_bot = (t * 0.25 + a) == a
    if np.any(no_bot & ~no_top):
        return None
    return np.where(no_top, np.float32(-np.inf), t32).astype(np.float32)


def height_bound(a, trees):
    """Per 16 x 16 cell, an FP32 upper bound of max(a, canopy top) with a the FP64 DSM (raised
    to >= 0) and top = trees + a (series.py:282-287): no sample of the sun-ray march in the cell
    can exceed it. Returned as the [3][3][R][C] table of its block maxima: level (i, j) holds at
    cell (r, c) the maximum over the cells [r, r + 2^i) x [c, c + 2^j) of the grid, so the
    maximum over any box of at most 8 x 8 cells is that of four blocks of one level."""
    a = np.asarray(a, dtype=np.float64)
    t = np.asarray(trees, dtype=np.float64)
    h = np.maximum(a, t + a)
    rows, cols = h.shape
    c = rt.HMAX_CELL
    R, C = -(-rows // c), -(-cols // c)
    pad = np.full((R * c, C * c), -np.inf)
    pad[:rows, :cols] = h
    m = pad.reshape(R, c, C, c).max(axis=(1, 3))
    m32 = m.astype(np.float32)
    low = m32.astype(np.float64) < m
    m32[low] = np.nextafter(m32[low], np.float32(np.inf))
    out = np.empty((3, 3, R, C), np.float32)
    for i in range(3):
        if i == 0:
            out[0, 0] = m32
        else:
            prev, k = out[i - 1, 0], 1 << (i - 1)
            out[i, 0] = prev
            out[i, 0, :R - k] = np.maximum(prev[:R - k], prev[k:])
        for j in range(1, 3):
            prev, k = out[i, j - 1], 1 << (j - 1)
            out[i, j] = prev
            out[i, j, :, :C - k] = np.maximum(prev[:, :C - k], prev[:, k:])
    return np.ascontiguousarray(out)


class PreparedDomain:
    """Device-resident static state of one raster domain."""

    def __init__(self, domain, *, stream=None, keep_host=False, trans=None):
        rows, cols = np.shape(domain.dsm)
        if rows * cols > rt.SW_MAX_PIXELS:
            raise ValueError(f'a {rows}x{cols} raster exceeds the {rt.SW_MAX_PIXELS}-pixel limit of one domain; '
                             f'split it into tiles (tiling.create_tiles)')
        self.domain = domain
        self.trans = float(CONSTANTS['trans'] if trans is None else trans)
        dsm = np.asarray(domain.dsm, dtype=np.float32)
        dem = np.asarray(domain.dem, dtype=np.float64)
        rows, cols = dsm.shape
        self.shape = (rows, cols)
        self.P = rows * cols
        a = dsm.astype(np.float64)
        raise_ = max(0.0, -float(a.min()))
        dem_raise = max(0.0, -float(dem.min()))
        if raise_:
            a = a + raise_
        if dem_raise:
            dem = dem + dem_raise
        self.dsm_raise = raise_
        trees = np.asarray(domain.trees, dtype=np.float32).copy()
        trees[trees < 0] = 0
        self.amaxvalue = float(max(a.max() - a.min(), float(trees.max())))
        ground = (a - dem) < 2.0
        z = np.zeros((rows, cols), dtype=np.float32)
        if domain.landcover is not None:
            lc = normalise_landcover(domain.landcover)
            maps = {k: np.copy(lc) for k in ('alb', 'emis', 'tgk', 'tstart', 'tmax')}
            for code in np.unique(lc):
                v = LANDCOVER[int(code)]
                sel = lc == code
                for key, val in zip(('alb', 'emis', 'tgk', 'tstart', 'tmax'), v):
                    maps[key][sel] = val
            w = LANDCOVER[99]
            self.landcover = 1
            self.wall_params = (float(np.float32(w[2])), float(np.float32(w[3])), float(np.float32(w[4])))
        else:
            maps = dict(alb=z + np.float32(CONSTANTS['albedo_g']), emis=z + np.float32(CONSTANTS['eground']),
                        tgk=z + np.float32(0.37), tstart=z + np.float32(-3.41), tmax=z + np.float32(15.))
            self.landcover = 0
            self.wall_params = (0.37, -3.41, 15.0)
        if domain.alb_grid is not None:
            maps['alb'] = np.asarray(domain.alb_grid, dtype=np.float32)
        if domain.emis_grid is not None:
            maps['emis'] = np.asarray(domain.emis_grid, dtype=np.float32)
        up = lambda x, t=np.float32: rt.DeviceArray.from_host(np.ascontiguousarray(x, dtype=t), stream=stream)
        self.dev = {}
        self.dev['dsm'] = up(dsm)
        self.dev['trees'] = up(trees)
        self.dev['walls'] = up(domain.walls)
        self.dev['aspect'] = up(domain.aspect)
        self.dev['ground'] = up(ground, np.uint8)
        for key in ('alb', 'emis', 'tgk', 'tstart', 'tmax'):
            self.dev[key] = up(maps[key])
        self.dev['hmax'] = up(height_bound(a, trees))
        mt = march_trees(a, trees)
        self.dev['march_trees'] = None if mt is None else up(mt)
        sv = sky_view(self.dev['dsm'], self.dev['trees'], raise_, self.shape, self.amaxvalue,
                      domain.scale, self.trans, stream, hmax=self.dev['hmax'], march_trees=self.dev['march_trees'])
        self.skyview = sv
        self.dev['vis'] = sv.bits
        # The derived planes are formed on the host with NumPy so that tan(asvf) - which decides
        # the sunlit/shaded wall split - is the same number the NumPy model uses.
        planes = sv.planes.to_host().reshape(16, rows, cols)
        sv.planes.free()                     # 128 B per pixel; the host copy is all that is read
        sv.planes = None
        svf = planes[0].astype(np.float32).astype(np.float64)
        svfveg = planes[5].astype(np.float32).astype(np.float64)
        self._svf, self._svfveg = svf, svfveg
        svfbuveg = svf - (1.0 - svfveg) * (1.0 - self.trans)
        asvf = np.arccos(np.sqrt(svf))
        tmp = svf + svfveg - 1.0
        tmp[tmp < 0.0] = 0.0
        with np.errstate(divide='ignore'):
            svfalfa = np.arcsin(np.exp(np.log(1.0 - tmp) / 2.0))
        self.dev['svfbuveg'] = up(svfbuveg)
        self.dev['svfveg'] = up(svfveg)
        self.dev['tan_asvf'] = up(np.tan(asvf), np.float64)
        self.dev['tan_svfalfa'] = up(np.tan(svfalfa), np.float64)
        self.dev['rec'] = rt.DeviceArray((self.P, 2), np.float32)
        self.dev['masks'] = rt.DeviceArray((self.P,), np.uint8)
        self.dev['lag'] = rt.DeviceArray((6, self.P), np.float32)
        self.dev['skysums'] = sky_sums(self.dev['vis'], self.P, stream)
        self._check_albedo(maps['alb'])
        ground_view_static(self, stream)
        if keep_host:
            self.host = dict(a=a, trees=trees, ground=ground, svf_planes=planes, svfbuveg=svfbuveg,
                             asvf=asvf, svfalfa=svfalfa, maps=maps)
        self.tile_cols = rt.TILE_X

    @property
    def tiles(self):
        """Workgroup tiles of this domain at its tile width."""
        return rt.tile_count(self.shape, self.tile_cols)

    def set_transmissivity(self, trans, stream=None):
        """Canopy transmissivity of the static svfbuveg plane (series.py:343)."""
        self.trans = float(trans)
        svfbuveg = self._svf - (1.0 - self._svfveg) * (1.0 - self.trans)
        self.dev['svfbuveg'].upload(svfbuveg, stream)

    def set_maps(self, alb=None, emis=None, stream=None):
        """Replace the ground albedo / emissivity maps (host arrays or device arrays); the static
        ground-view albedo sums follow the new albedo map."""
        for key, v in (('alb', alb), ('emis', emis)):
            if v is None:
                continue
            if isinstance(v, (rt.DeviceArray, rt.DeviceView)):
                rt.call('sw_memcpy_d2d', self.dev[key].ptr, v.ptr, self.dev[key].nbytes, rt._s(stream))
            else:
                if key == 'alb':
                    self._check_albedo(v)
                self.dev[key].upload(np.asarray(v, dtype=np.float32).reshape(self.shape), stream)
        if alb is not None:
            ground_view_static(self, stream)

    @staticmethod
    def _check_albedo(alb):
        """The hour record stores albedo x shadow with its sign bit free for the sunlit-wall flag."""
        a = np.asarray(alb, dtype=np.float32)
        if not np.all(a >= 0):
            raise ValueError('albedo maps must be >= 0')

    def struct(self, **extra):
        """The sw_domain record for this domain (extra: hours, ptab_f, ... device pointers)."""
        d = self.dev
        s = rt.SwDomain()
        s.rows, s.cols, s.pixels, s.landcover = self.shape[0], self.shape[1], self.P, self.landcover
        s.dsm_raise = self.dsm_raise
        s.tile_cols = self.tile_cols
        for field, key in (('dsm', 'dsm'), ('trees', 'trees'), ('walls', 'walls'), ('aspect', 'aspect'),
                           ('emis', 'emis'), ('alb', 'alb'), ('tgk', 'tgk'), ('tstart', 'tstart'),
                           ('tmaxlst', 'tmax'), ('svfbuveg', 'svfbuveg'), ('tan_asvf', 'tan_asvf'),
                           ('tan_svfalfa', 'tan_svfalfa'), ('ground', 'ground'), ('vis', 'vis'),
                           ('rec', 'rec'), ('lag', 'lag'), ('skysums', 'skysums'), ('hmax', 'hmax'),
                           ('masks', 'masks'), ('gvf_len', 'gvf_len'), ('gvf_nosh', 'gvf_nosh'),
                           ('gvf_edge', 'gvf_edge')):
            setattr(s, field, d[key].ptr)
        s.march_trees = d['march_trees'].ptr if d.get('march_trees') is not None else 0
        s.hmax_cols = -(-self.shape[1] // rt.HMAX_CELL)
        for k, v in extra.items():
            setattr(s, k, v or 0)
        return s


def ground_view_static(prep, stream=None):
    """(Re)compute the static ground-view planes of a prepared domain (sw_gvf_static): per
    direction the steps before the first non-ground sample (bit 7: the wall faces away from the
    direction), and the unshadowed-albedo sums."""
    from .sky import gvf_geometry
    off, bounds, first, second = gvf_geometry(prep.domain.scale)
    d = prep.dev
    if 'gvf_len' not in d:
        d['gvf_off'] = rt.DeviceArray.from_host(off, np.int32)
        d['gvf_bounds'] = rt.DeviceArray.from_host(bounds, np.float64)
        d['gvf_len'] = rt.DeviceArray((5, prep.P), np.uint32)
        d['gvf_nosh'] = rt.DeviceArray((10, prep.P), np.float32)
        d['gvf_edge'] = rt.DeviceArray((5, prep.P), np.uint32)
    rows, cols = prep.shape
    rt.call('sw_gvf_static', d['ground'].ptr, d['alb'].ptr, d['walls'].ptr, d['aspect'].ptr, rows, cols,
            d['gvf_off'].ptr, d['gvf_bounds'].ptr, first, second, d['gvf_len'].ptr, d['gvf_nosh'].ptr,
            d['gvf_edge'].ptr, rt._s(stream))
    rt.call('sw_stream_sync', rt._s(stream))


def pack_visibility(shmat, vegshmat, vbshvegshmat):
    """[rows, cols, 153] 0/1 masks -> uint32 [3, 5, P] bit words (bit k of word w = patch 32w+k)."""
    out = []
    for m in (shmat, vegshmat, vbshvegshmat):
        m = np.asarray(m)
        rows, cols, n = m.shape
        bits = np.zeros((rows * cols, rt.SW_VIS_WORDS * 32), dtype=np.uint8)
        bits[:, :n] = (m.reshape(rows * cols, n) > 0.5)
        words = np.packbits(bits, axis=1, bitorder='little').view('<u4')
        out.append(words.T)
    return np.ascontiguousarray(np.stack(out), dtype=np.uint32)


def prepared_from_arrays(*, dsm, vegdem, walls, dirwalls, buildings, svfbuveg, asvf, svfalfa, shmat,
                         vegshmat, vbshvegshmat, alb_grid, emis_grid, TgK, Tstart, TmaxLST, landcover,
                         wall_params, amaxvalue, scale=1.0, stream=None):
    """A PreparedDomain from the static dict series._prepare_spatial builds (series.py:350-365),
    without recomputing the sky view. DSM and canopy heights must be FP32-exact."""
    a = np.asarray(dsm, dtype=np.float64)
    top = np.asarray(vegdem, dtype=np.float64)
    rows, cols = a.shape
    dsm32 = a.astype(np.float32)
    trees32 = np.where(top != 0, top - a, 0.0).astype(np.float32)
    if not np.array_equal(dsm32.astype(np.float64), a):
        raise ValueError('dsm must be FP32-exact')
    if np.any(trees32 < 0):
        raise ValueError('vegdem below the DSM: canopy heights must be >= 0 (series.py:282-287)')
    self = PreparedDomain.__new__(PreparedDomain)
    from types import SimpleNamespace
    self.domain = SimpleNamespace(scale=float(scale), dsm=dsm32, trees=trees32, walls=walls, aspect=dirwalls,
                                  latitude=None, met=None)
    self.shape, self.P = (rows, cols), rows * cols
    self.dsm_raise = 0.0
    self.amaxvalue = float(amaxvalue)
    self.landcover = int(landcover)
    self.wall_params = tuple(float(v) for v in wall_params)
    plane = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float32), (rows, cols))
    up = lambda x, t=np.float32: rt.DeviceArray.from_host(np.ascontiguousarray(x, dtype=t), stream=stream)
    self.dev = dict(dsm=up(dsm32), trees=up(trees32), walls=up(plane(walls)), aspect=up(plane(dirwalls)),
                    ground=up(np.asarray(buildings) == 1, np.uint8), alb=up(plane(alb_grid)),
                    emis=up(plane(emis_grid)), tgk=up(plane(TgK)), tstart=up(plane(Tstart)),
                    tmax=up(plane(TmaxLST)), svfbuveg=up(plane(svfbuveg)),
                    tan_asvf=up(np.tan(np.asarray(asvf, dtype=np.float64)), np.float64),
                    tan_svfalfa=up(np.tan(np.asarray(svfalfa, dtype=np.float64)), np.float64),
                    vis=up(pack_visibility(shmat, vegshmat, vbshvegshmat), np.uint32),
                    rec=rt.DeviceArray((rows * cols, 2), np.float32),
                    masks=rt.DeviceArray((rows * cols,), np.uint8),
                    lag=rt.DeviceArray((6, rows * cols), np.float32))
    self.dev['skysums'] = sky_sums(self.dev['vis'], rows * cols, stream)
    self.dev['hmax'] = up(height_bound(a, trees32))
    mt = march_trees(a, trees32)
    self.dev['march_trees'] = None if mt is None else up(mt)
    PreparedDomain._check_albedo(plane(alb_grid))
    ground_view_static(self, stream)
    self.skyview = None
    self.tile_cols = rt.TILE_X
    return self


def model_inputs(prep, location=None):
    """The static dict of series._prepare_spatial (series.py:343-365) for a domain prepared with
    ``PreparedDomain(domain, keep_host=True)``: the keyword inputs of model.prepare_model and the
    static arguments of Solweig_2022a_calc. SVF planes are rounded to float32 (series.py:337) in
    the UMEP tuple order, masks are uint8 [rows, cols, 153] unpacked from the device bits, diffsh
    is compute_diffsh's FP32 3-value plane (svf.py:117-131)."""
    from .svf import UMEP_NAMES, UMEP_ORDER, compute_diffsh, unpack_visibility
    h = prep.host
    a = h['a']
    t = np.asarray(h['trees'], dtype=np.float64)
    rows, cols = prep.shape
    vegdem = t + a
    vegdem[vegdem == a] = 0
    vegdem2 = t * 0.25 + a
    vegdem2[vegdem2 == a] = 0
    bush = np.logical_not(vegdem2 * vegdem) * vegdem
    planes = np.asarray(h['svf_planes'])
    svf = {name: planes[k].astype(np.float32).astype(np.float64) for name, k in zip(UMEP_NAMES, UMEP_ORDER)}
    shm, vgm, vbm = unpack_visibility(prep.dev['vis'].to_host(), prep.shape)
    d = prep.domain
    maps = h['maps']
    f64 = lambda x: np.asarray(x, dtype=np.float32).astype(np.float64)
    if prep.landcover:
        lc = normalise_landcover(d.landcover).astype(np.float64)
    else:
        lc = np.zeros((rows, cols))
    loc = location or {'latitude': float(d.latitude), 'longitude': float(d.longitude),
                       'altitude': float(getattr(d, 'site_altitude', 3.0))}
    out = dict(dsm=a, scale=float(d.scale), rows=rows, cols=cols, vegdem=vegdem, vegdem2=vegdem2,
               albedo_b=CONSTANTS['albedo_b'], absK=CONSTANTS['absK'], absL=CONSTANTS['absL'],
               ewall=CONSTANTS['ewall'], Fside=CONSTANTS['Fside'], Fup=CONSTANTS['Fup'], Fcyl=CONSTANTS['Fcyl'],
               usevegdem=1, onlyglobal=1, buildings=np.where(h['ground'], 1.0, 0.0), location=loc,
               landcover=int(prep.landcover), lc_grid=lc, dirwalls=f64(d.aspect), walls=f64(d.walls), cyl=1,
               elvis=0.0, amaxvalue=float(prep.amaxvalue), bush=bush, TgK=f64(maps['tgk']), Tstart=f64(maps['tstart']),
               alb_grid=f64(maps['alb']), emis_grid=f64(maps['emis']), TgK_wall=prep.wall_params[0],
               Tstart_wall=prep.wall_params[1], TmaxLST=f64(maps['tmax']), TmaxLST_wall=prep.wall_params[2],
               first=1.0, second=22.0, svfalfa=h['svfalfa'], svfbuveg=h['svfbuveg'],
               diffsh=compute_diffsh(shm, vgm, prep.trans), shmat=shm.astype(np.uint8), vegshmat=vgm.astype(np.uint8),
               vbshvegshmat=vbm.astype(np.uint8), anisotropic_sky=1, asvf=h['asvf'], patch_option=2, **svf)
    return out
