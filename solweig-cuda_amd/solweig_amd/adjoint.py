"""Gradient of J = mean over ground pixels of UTCI at one hour of a day (hand-written adjoint).

Parameters: the per-pixel ground albedo and emissivity maps, the wall albedo ``albedo_b``, the
canopy transmissivity ``trans`` (both the leaf-on shadow transmissivity psi of series.py:416-422
and transVeg of svfbuveg / diffsh, series.py:343-349), and the hourly Ta, RH, wind and global
radiation of the forcing (columns 11, 10, 9, 14). Shadow and sky-view masks are held fixed.

The per-pixel chain runs on the device (``sw_day_adjoint``): the forward of the hours before
the objective hour, then a backward sweep from the objective hour to the day's first daytime
hour that recomputes each hour's shadow pass and ground-view walk instead of storing them. The
backward needs one forward state, the lag planes entering the objective hour (the lag recursion
is linear, so earlier hours need none): ``objective()`` keeps a copy of them, and a gradient of
the same parameters starts from that copy instead of running the forward again. It
returns dJ/d(albedo map), dJ/d(emissivity map), per hour dJ/d(value) for the hour-record values
the kernels read (Ta, psi, CI_TgG, wall flux, wall albedo, radI, radD, radG, sigma T^4, the wall
and sky flux factors, the UTCI coefficients) and, at the objective hour, per-patch sums for the
luminance and diffuse tables. The hour-record values are closed-form scalar functions of the
forcing computed on the host in FP64 (sky.hour_tables: clearness index, Reindl split, Perez
luminance, Prata sky emissivity, the UTCI coefficient polynomial); their derivatives with
respect to the forcing are taken by central differences of that FP64 builder, then combined
with the device's dJ/d(value).

The objective is the reference's UTCI map at one step of series._run_day (series.py:445-469) with
outputs NaN off the ground mask; the adjoint has no counterpart in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import runtime as rt
from .batched import Engine
from .sky import CONSTANTS, hour_branches, hour_tables
from .static import PreparedDomain
from .sun import met_sun_device

# hour-record values whose derivative the device accumulates (solweig_hip.h SWA_* order)
VALUES = ('TA', 'PSI', 'CITGG', 'LWALL', 'ALBEDO_B', 'RADI', 'RADD', 'RADG', 'FLUX4', 'SW_SUNLIT',
          'SW_SHADED', 'SKY_A', 'SKY_B', 'LW_AMBIENT', 'LW_SUNLIT', 'TA32')
FORCING = {'Ta': 11, 'RH': 10, 'Ws': 9, 'radG': 14}
STEPS = {'Ta': 1e-4, 'RH': 1e-3, 'Ws': 1e-5, 'radG': 1e-3, 'albedo_b': 1e-6, 'trans': 1e-6}


class UtciGradient:
    """One domain, one day of forcing, one objective hour (default 14:00 row)."""

    def __init__(self, domain, met=None, objective_hour=14, albedo_b=CONSTANTS['albedo_b'],
                 trans=CONSTANTS['trans'], stream=None):
        if getattr(domain, 'landcover', None) is not None:
            raise ValueError('the adjoint supports the default land cover (no land-cover grid)')
        self.domain = domain
        self.met = np.array(domain.met if met is None else met, dtype=np.float64)
        self.n_hours = self.met.shape[0]
        self.h_obj = int(objective_hour)
        if not 0 <= self.h_obj < self.n_hours:
            raise ValueError('objective hour outside the forcing rows')
        self.albedo_b, self.trans = float(albedo_b), float(trans)
        self._cached = None
        self._maps_version = 0
        self._ckpt_key = None
        self.stream = stream or rt.Stream()
        self.pd = PreparedDomain(domain, stream=self.stream, keep_host=True, trans=self.trans)
        self.eng = Engine([self.pd], n_hours=self.n_hours, stream=self.stream)
        ground = self.pd.host['ground']
        self.count = int(np.count_nonzero(ground))
        if self.count == 0:
            raise ValueError('no ground pixels')
        P = self.pd.P
        rows, cols = self.pd.shape
        # workgroup partial sums: one column per tile of the widest-tile count any width gives
        self.partial_cols = max(-(-rows // (256 // w)) * -(-cols // w) for w in rt.TILE_CHOICES)
        self.buf = dict(lam_state=rt.DeviceArray((5, P), np.float32), lam_rec=rt.DeviceArray((3, P), np.float32),
                        coef=rt.DeviceArray((18, P, 4), np.float32), grad_alb=rt.DeviceArray((P,), np.float32),
                        grad_emis=rt.DeviceArray((P,), np.float32), lam_patch=rt.DeviceArray((2, P), np.float32),
                        scalars=rt.DeviceArray((self.n_hours, rt.ADJ_SLOTS), np.float64),
                        patch=rt.DeviceArray((4, rt.SW_PATCHES), np.float64),
                        objective=rt.DeviceArray((2,), np.float64),
                        partial=rt.DeviceArray((rt.ADJ_PARTIAL_SLOTS, self.partial_cols), np.float64))
        # lag planes entering the objective hour, kept by objective() for the gradient
        self._ckpt = rt.DeviceArray(self.pd.dev['lag'].shape, np.float32)
        a = rt.SwAdjoint()
        for k, v in self.buf.items():
            setattr(a, k, v.ptr)
        a.svfveg = self.pd.dev['svfveg'].ptr
        self._adj_host = a
        self.adj_dev = rt.DeviceArray((ctypes.sizeof(a),), np.uint8)
        rt.call('sw_memcpy_h2d', self.adj_dev.ptr, ctypes.addressof(a), ctypes.sizeof(a), self.stream.handle)
        self.stream.sync()

    # ------------------------------------------------------------------ parameters
    def set_parameters(self, alb=None, emis=None, albedo_b=None, trans=None, met=None):
        self.pd.set_maps(alb=alb, emis=emis, stream=self.stream)
        if alb is not None or emis is not None:
            self._maps_version += 1
        if albedo_b is not None:
            self.albedo_b = float(albedo_b)
        if trans is not None and float(trans) != self.pd.trans:
            self.trans = float(trans)
            self.pd.set_transmissivity(self.trans, self.stream)
        if met is not None:
            self.met = np.array(met, dtype=np.float64)
        self.stream.sync()

    def _constants(self, albedo_b=None, trans=None):
        c = dict(CONSTANTS)
        c['albedo_b'] = self.albedo_b if albedo_b is None else albedo_b
        c['trans'] = self.trans if trans is None else trans
        return c

    def clear_cache(self):
        """Forget the hour tables and the objective's lag planes (the next objective or gradient
        rebuilds them, as after a change of parameters)."""
        self._cached = None
        self._ckpt_key = None

    def _tables(self):
        """Sun positions and hour tables of the current forcing and scalar parameters, built once
        per parameter set: the gradient after an objective of the same parameters reuses the
        objective's tables (they are host data of a few hundred KB, not per-pixel state)."""
        key = (self.albedo_b, self.trans, self.met.tobytes())
        if self._cached is None or self._cached[0] != key:
            sun = met_sun_device(self.met, self.domain.location, self.domain.utc)
            tabs = self.eng.tables(met=[self.met], sun=[sun], constants=self._constants(), psi=self.trans)
            self._cached = (key, sun, tabs)
        return self._cached[1], self._cached[2]

    # ------------------------------------------------------------------ forward
    def objective(self):
        """J from the batched forward (hours 0 .. objective hour)."""
        sun, tabs = self._tables()
        self.eng.load(tabs)
        self.eng.reset_state()
        self.eng.diffuse()
        lag = self.pd.dev['lag']
        for h0, h1 in ((0, self.h_obj), (self.h_obj, self.h_obj + 1)):
            if h0 == self.h_obj:
                rt.call('sw_memcpy_d2d', self._ckpt.ptr, lag.ptr, lag.nbytes, self.stream.handle)
            if h1 > h0:
                rt.call('sw_hours_forward', self.eng.dom_dev.ptr, self.eng.block_start.ptr, 1,
                        self.eng.total_blocks, h0, h1, self.eng.gvf_off.ptr, self.eng.gvf_bounds.ptr,
                        self.eng.first, self.eng.second, self.stream.handle)
        self._ckpt_key = (self._cached[0], self._maps_version)
        u = self.eng.buf[0]['utci'].plane(self.h_obj).to_host(self.stream)
        self.stream.sync()
        return float(np.nanmean(u))

    # ------------------------------------------------------------------ gradient
    def run_device(self):
        """Queue the forward + backward sweep (the device part of the gradient)."""
        sun, tabs = self._tables()
        t = tabs[0]
        if not t.hours[self.h_obj, rt.H['DAY']]:
            raise ValueError('the objective hour must be a daytime hour')
        first = self.h_obj
        while first > 0 and t.hours[first - 1, rt.H['DAY']] and not t.hours[first, rt.H['LAG_FIRST']]:
            first -= 1
        self._last = (sun, t, first)
        if self.eng.total_blocks > self.partial_cols:
            raise ValueError('more tiles than the partial-sum buffer holds')
        self.eng.load(tabs)
        self.eng.reset_state()
        for v in self.buf.values():
            v.zero(self.stream)
        e = self.eng
        if self._ckpt_key == (self._cached[0], self._maps_version):
            # the objective of these parameters ran last: start the backward from its lag planes
            lag = self.pd.dev['lag']
            rt.call('sw_memcpy_d2d', lag.ptr, self._ckpt.ptr, lag.nbytes, self.stream.handle)
            for h in range(self.h_obj, first - 1, -1):
                rt.call('sw_hour_shadow', e.dom_dev.ptr, e.block_start.ptr, 1, e.total_blocks, h, 0,
                        self.stream.handle)
                rt.call('sw_adjoint_hour', e.dom_dev.ptr, e.block_start.ptr, e.total_blocks, h, self.h_obj,
                        self.adj_dev.ptr, e.gvf_off.ptr, e.gvf_bounds.ptr, e.first, e.second, 1.0 / self.count,
                        self.stream.handle)
        else:
            rt.call('sw_day_adjoint', e.dom_dev.ptr, e.block_start.ptr, e.total_blocks, self.h_obj, first,
                    self.adj_dev.ptr, e.gvf_off.ptr, e.gvf_bounds.ptr, e.first, e.second, 1.0 / self.count,
                    self.stream.handle)

    def gradient(self):
        """J and dJ/d every parameter (host arrays / floats). The derivatives of the hour tables
        are taken on the host while the device runs the sweep."""
        self.run_device()
        sun, t, first = self._last
        jac = {h: self._jacobian(sun, t, h, h == self.h_obj) for h in range(first, self.h_obj + 1)}
        self.stream.sync()
        obj = self.buf['objective'].to_host()
        scal = self.buf['scalars'].to_host()
        patch = self.buf['patch'].to_host()
        g_alb = self.buf['grad_alb'].to_host().reshape(self.pd.shape)
        g_emis = self.buf['grad_emis'].to_host().reshape(self.pd.shape)
        out = dict(J=float(obj[0] / self.count), ground_pixels=int(obj[1]), alb=g_alb, emis=g_emis,
                   hours=(first, self.h_obj))
        for name in FORCING:
            out[name] = np.zeros(self.n_hours)
        out['albedo_b'] = 0.0
        out['trans'] = float(scal[:, rt.ADJ['TRANS_STATIC']].sum())
        for h in range(first, self.h_obj + 1):
            dv = self._value_gradient(scal[h], patch if h == self.h_obj else None, t, h)
            for name, col in jac[h].items():
                if name in FORCING:
                    out[name][h] += float(dv @ col)
                else:
                    out[name] += float(dv @ col)
        return out

    def _jacobian(self, sun, t, h, top):
        """d(value vector)/d(parameter) of hour h. Ta, RH, radG (and Ws at the objective hour,
        the only hour whose UTCI counts) by central differences of the FP64 table builder; the
        wall albedo and the transmissivity enter the values linearly and are written out."""
        row = self.met[h]
        base = dict(Ta=float(row[11]), RH=float(row[10]), Ws=float(row[9]), radG=float(row[14]),
                    albedo_b=self.albedo_b, trans=self.trans)
        names = ['Ta', 'RH', 'radG'] + (['Ws'] if top else [])
        ci_in = t.meta[h]['state_in'][0]

        def vec(p):
            return self._vector(sun, t, h, p['Ta'], p['RH'], p['Ws'], p['radG'], p['albedo_b'], p['trans'], top)

        def branches(p):
            return hour_branches(float(sun['altitude'][h]), float(sun['azimuth'][h]), float(sun['zen'][h]),
                                 float(sun['jday'][h]), p['Ta'], p['RH'], p['radG'], float(row[12]),
                                 self.domain.latitude, ci_in)

        b0, v0 = branches(base), None
        cols = {}
        for name in names:
            # central differences inside the hour's branch (Reindl and Perez bins, CI clamps);
            # one-sided on the side that stays in it when the forcing sits within a step of an edge
            step = STEPS[name]
            for _ in range(3):
                hi, lo = dict(base), dict(base)
                hi[name] += step
                lo[name] -= step
                same_hi, same_lo = branches(hi) == b0, branches(lo) == b0
                if same_hi and same_lo:
                    cols[name] = (vec(hi) - vec(lo)) / (2.0 * step)
                elif same_hi or same_lo:
                    v0 = vec(base) if v0 is None else v0
                    cols[name] = (vec(hi) - v0) / step if same_hi else (v0 - vec(lo)) / step
                else:
                    step *= 0.01
                    continue
                break
            else:
                raise RuntimeError(f'hour {h}: the forcing sits on a branch edge of {name}')
        n = len(VALUES) + 7 + (2 * rt.SW_PATCHES + 1 if top else 0)
        hh = t.hours[h]
        ab, tr = np.zeros(n), np.zeros(n)
        ab[VALUES.index('ALBEDO_B')] = 1.0
        ab[VALUES.index('SW_SUNLIT')] = hh[rt.H['RADI']] * hh[rt.H['COSALT']] / np.pi
        ab[VALUES.index('SW_SHADED')] = hh[rt.H['RADD']] * 0.5 / np.pi
        tr[VALUES.index('PSI')] = 1.0
        if top:
            tr[-1] = 1.0                                   # DIFF_VEG = 1 - (1 - trans)
        cols['albedo_b'], cols['trans'] = ab, tr
        return cols

    # ------------------------------------------------------------------ host chain
    def _vector(self, sun, t, h, Ta, RH, Ws, radG, albedo_b, trans, with_patches):
        """The hour-record values the device differentiates, as smooth FP64 functions."""
        row = self.met[h]
        m = t.meta[h]
        hh, pf, _, _, _, _, meta = hour_tables(
            float(sun['altitude'][h]), float(sun['azimuth'][h]), float(sun['zen'][h]), float(sun['jday'][h]),
            float(sun['dectime'][h]), float(sun['altmax'][h]), Ta, RH, radG, float(row[12]), Ws, row[1],
            self.domain.latitude, self.pd.shape, self.domain.scale, -1.0, m['state_in'],
            wall_params=self.pd.wall_params, constants=self._constants(albedo_b, trans), psi=trans, f64=True,
            rows=with_patches, ray=False)
        v = [hh[rt.H[k]] for k in VALUES] + list(hh[rt.H['UTCI0']:rt.H['UTCI0'] + 7])
        if with_patches:
            v += list(pf[:, rt.PF['LV']]) + list(pf[:, rt.PF['KD']]) + [hh[rt.H['DIFF_VEG']]]
        return np.asarray(v, dtype=np.float64)

    def _value_gradient(self, slots, patch, t, h):
        """dJ/d(value vector) of hour h, in the order of _vector."""
        g = [slots[rt.ADJ[k]] for k in VALUES] + list(slots[rt.ADJ['UTCI0']:rt.ADJ['UTCI0'] + 7])
        if patch is not None:
            hh = t.hours[h]
            radD, dveg = hh[rt.H['RADD']], hh[rt.H['DIFF_VEG']]
            lv = t.ptab_f[h, :, rt.PF['LV']].astype(np.float64)
            kd = t.ptab_f[h, :, rt.PF['KD']].astype(np.float64)
            P = rt.ADJ_PATCH
            g_lv = radD * (patch[P['SKY_KDOWN']] + dveg * patch[P['VEG_KDOWN']])
            g_kd = patch[P['SKY_KSIDE']] + dveg * patch[P['VEG_KSIDE']]
            g_dv = radD * float(lv @ patch[P['VEG_KDOWN']]) + float(kd @ patch[P['VEG_KSIDE']])
            g += list(g_lv) + list(g_kd) + [g_dv]
        return np.asarray(g, dtype=np.float64)
