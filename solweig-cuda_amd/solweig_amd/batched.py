"""Batched day forward: every hour of a day, for one or many uneven domains, on one stream.

The solar position of every forcing row comes from the device (FP64 SPA, sw_sun_position, once
per domain for its whole forcing). Per day the host builds each domain's hour tables
(sky.build_day) and uploads them into buffers of fixed capacity; the device then runs the day
either resident (sw_day_forward: the shadow pass of every hour in one launch, then one pixel pass
in which each tile runs through the hours with its lag state in registers) or as one shadow and
one pixel launch per hour (sw_hours_forward), each launch covering the tiles of all domains
together. Lag state is cleared at each day start (series.py:425-436). Because buffer addresses
never change, a day's launches can be captured once as a hipGraph and replayed for every further
day after the tables are refreshed.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes
import multiprocessing as mp
import os

import numpy as np

from . import runtime as rt
from .sky import CONSTANTS, STEP_BLOCK, build_day, gvf_geometry
from .static import PreparedDomain
from .sun import met_sun, met_sun_device


class Engine:
    def __init__(self, domains, n_hours=24, stream=None, keep_outputs=True, extra_outputs=False, resident=True):
        """``extra_outputs``: also keep each hour's Kdown, Kup, Ldown, Lup and Shadow planes
        ([n_hours][5][P] per domain, rt.EXTRA_NAMES order), the optional series outputs.
        ``resident``: run a day as two launches (sw_day_forward: the shadow pass of every hour into
        per-hour record and mask planes, 9 B per pixel-hour, then one pixel pass in which each tile
        runs through the hours with its lag state in registers) instead of two launches per hour
        (sw_hours_forward)."""
        scales = {float(d.domain.scale if isinstance(d, PreparedDomain) else d.scale) for d in domains}
        if len(scales) != 1:
            # one launch shares one ground-view step table (the step shifts and counts follow the
            # pixel size, umep_solweig.py:1466-1476)
            raise ValueError(f'the domains of one engine must share one pixel size, got {sorted(scales)}; '
                             f'run each pixel size in an engine of its own')
        self.domains = [d if isinstance(d, PreparedDomain) else PreparedDomain(d) for d in domains]
        self.n_hours = int(n_hours)
        self.stream = stream if stream is not None else rt.Stream()
        off, bounds, self.first, self.second = gvf_geometry(self.domains[0].domain.scale)
        self.gvf_off = rt.DeviceArray.from_host(off, np.int32)
        self.gvf_bounds = rt.DeviceArray.from_host(bounds, np.float64)
        self.buf = []
        for d in self.domains:
            cap = self.n_hours * (max(d.shape) + 2 + STEP_BLOCK)
            b = dict(hours=rt.DeviceArray((self.n_hours, rt.SW_HOUR_STRIDE), np.float64),
                     ptab_f=rt.DeviceArray((self.n_hours, rt.SW_PATCHES, rt.SW_PATCH_F), np.float32),
                     ptab_d=rt.DeviceArray((self.n_hours, rt.SW_PATCH_D, rt.SW_PATCHES), np.float64),
                     ptab_s=rt.DeviceArray((self.n_hours, rt.SW_SEARCH_WORDS), np.uint32),
                     ray_off=rt.DeviceArray((cap, 2), np.int32), ray_dz=rt.DeviceArray((cap,), np.float64),
                     cap=cap)
            if keep_outputs:
                b['tmrt'] = rt.DeviceArray((self.n_hours, d.P), np.float32)
                b['utci'] = rt.DeviceArray((self.n_hours, d.P), np.float32)
            if extra_outputs:
                b['extra'] = rt.DeviceArray((self.n_hours, rt.SW_EXTRA_PLANES, d.P), np.float32)
            # the shadow pass's record (Lup, albedo x shadow | sunlit wall) and mask byte of every
            # hour; the single-hour launches use the first slot
            b['rec'] = rt.DeviceArray((self.n_hours if resident else 1, d.P, 2), np.float32)
            b['masks'] = rt.DeviceArray((self.n_hours if resident else 1, d.P), np.uint8)
            self.buf.append(b)
        self.resident = bool(resident)
        self._sun_days = [None] * len(self.domains)
        self.block_start = rt.DeviceArray((len(self.domains) + 1,), np.int32)
        self.structs = (rt.SwDomain * len(self.domains))()
        self.dom_dev = rt.DeviceArray((ctypes.sizeof(self.structs),), np.uint8)
        self._write_domains()
        # two page-locked staging slots per domain: the host fills one while the copies from the
        # other may still be queued behind the previous day's launches
        self._stage = [[dict(hours=rt.PinnedArray(b['hours'].shape, np.float64),
                             ptab_f=rt.PinnedArray(b['ptab_f'].shape, np.float32),
                             ptab_d=rt.PinnedArray(b['ptab_d'].shape, np.float64),
                             ptab_s=rt.PinnedArray(b['ptab_s'].shape, np.uint32),
                             ray_off=rt.PinnedArray((b['cap'], 2), np.int32),
                             ray_dz=rt.PinnedArray((b['cap'],), np.float64)) for b in self.buf]
                       for _ in range(2)]
        self._stage_event = [rt.Event(), rt.Event()]
        self._stage_used = [False, False]
        self._slot = 0
        self.graph = None

    def _write_domains(self):
        """Upload the domain records and the tile prefix (after a change of tile widths)."""
        starts = np.zeros(len(self.domains) + 1, dtype=np.int32)
        starts[1:] = np.cumsum([d.tiles for d in self.domains])
        self.total_blocks = int(starts[-1])
        self.block_start.upload(starts, self.stream)
        for i, (d, b) in enumerate(zip(self.domains, self.buf)):
            self.structs[i] = d.struct(hours=b['hours'].ptr, ptab_f=b['ptab_f'].ptr, ptab_d=b['ptab_d'].ptr,
                                       ptab_s=b['ptab_s'].ptr, ray_off=b['ray_off'].ptr, ray_dz=b['ray_dz'].ptr,
                                       tmrt=b['tmrt'].ptr if 'tmrt' in b else 0,
                                       utci=b['utci'].ptr if 'utci' in b else 0,
                                       full=b['extra'].ptr if 'extra' in b else 0,
                                       rec=b['rec'].ptr, masks=b['masks'].ptr)
        rt.call('sw_memcpy_h2d', self.dom_dev.ptr, ctypes.addressof(self.structs),
                ctypes.sizeof(self.structs), self.stream.handle)
        self.stream.sync()
        self.graph = None                                # launch geometry changed

    def set_tiles(self, tile_cols):
        """Tile width per domain (one value for all, or a list), from rt.TILE_CHOICES."""
        widths = [tile_cols] * len(self.domains) if np.ndim(tile_cols) == 0 else list(tile_cols)
        for d, w in zip(self.domains, widths):
            if int(w) not in rt.TILE_CHOICES:
                raise ValueError(f'tile width {w} not in {rt.TILE_CHOICES}')
            d.tile_cols = int(w)
        self._write_domains()

    # ------------------------------------------------------------------ tables
    def _sun(self, k, day_index, rows):
        """Solar position of one day's rows of domain k, on the device (met_sun_device): the
        domain's whole forcing is evaluated once, in chunks of n_hours rows (one chunk per day, as
        a per-day met_sun call takes it), and sliced per day; other rows are evaluated on call."""
        dm = self.domains[k].domain
        if rows is not None:
            return met_sun_device(rows, dm.location, dm.utc)
        if self._sun_days[k] is None:
            self._sun_days[k] = met_sun_device(dm.met, dm.location, dm.utc, chunk_rows=self.n_hours,
                                               stream=self.stream)
        lo, hi = day_index * self.n_hours, (day_index + 1) * self.n_hours
        return {key: v[lo:hi] for key, v in self._sun_days[k].items()}

    def _jobs(self, day_index, met, sun, constants, psi):
        for k, d in enumerate(self.domains):
            dm = d.domain
            rows = met[k] if met is not None else dm.met[day_index * self.n_hours:(day_index + 1) * self.n_hours]
            s = sun[k] if sun is not None else self._sun(k, day_index, None if met is None else np.asarray(rows))
            yield (np.asarray(rows), s, dm.location, dm.utc, dm.latitude, d.shape,
                   dm.scale, d.amaxvalue, d.landcover, d.wall_params, constants or CONSTANTS, psi)

    def tables(self, day_index=0, met=None, sun=None, constants=None, psi=None):
        """Host tables of one day for every domain. ``met`` / ``sun`` override per domain;
        ``constants`` (sky.CONSTANTS-like) and ``psi`` (canopy transmissivity of the hour's
        shadow, default by season) override the model constants."""
        return [_day_tables(*job) for job in self._jobs(day_index, met, sun, constants, psi)]

    def start_table_workers(self, workers=None):
        """Build day tables in ``workers`` CPU processes from now on (``tables_async``). The
        tables are per-hour host FP64 work (clearness, Perez, patch rows, search tables, rays);
        one day of one domain is one job, so days ahead and domains spread over
        the host's cores while the device runs. Workers are spawned and touch no GPU."""
        if workers is None:
            workers = max(1, min(8, (len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity')
                                     else os.cpu_count() or 2) - 1))
        self.close_table_workers()
        self._pool = cf.ProcessPoolExecutor(int(workers), mp_context=mp.get_context('spawn'))
        # import the table code in every worker now, not inside the first timed day
        list(self._pool.map(_warm, range(int(workers))))
        return int(workers)

    def close_table_workers(self):
        pool, self._pool = getattr(self, '_pool', None), None
        if pool is not None:
            pool.shutdown(wait=True, cancel_futures=True)

    def tables_async(self, day_index=0, met=None, sun=None, constants=None, psi=None):
        """``tables`` on the table workers: returns an object whose ``result()`` is the list of
        DayTables. Without workers (start_table_workers not called) the tables are built in this
        process when ``result()`` is called."""
        if getattr(self, '_pool', None) is None:
            jobs = list(self._jobs(day_index, met, sun, constants, psi))
            return _Lazy(lambda: [_day_tables(*job) for job in jobs])
        return _Pending([self._pool.submit(_day_tables, *job)
                         for job in self._jobs(day_index, met, sun, constants, psi)])

    def load(self, tables):
        """Queue the upload of one day's tables on the engine's stream (fixed device addresses, so
        a captured graph stays valid). The copies run after every launch queued before them, so
        the host may build and load the next day while the device still runs this one."""
        for t, b in zip(tables, self.buf):
            if t.n_hours != self.n_hours:
                raise ValueError(f'expected {self.n_hours} hours, got {t.n_hours}')
            if t.ray_dz.shape[0] > b['cap']:
                raise ValueError('ray tables exceed their buffer')
        slot = self._slot
        self._slot ^= 1
        if self._stage_used[slot]:
            self._stage_event[slot].sync()        # the copies out of this slot have finished
        h = self.stream.handle
        for t, b, st in zip(tables, self.buf, self._stage[slot]):
            n = t.ray_dz.shape[0]
            st['hours'].array[...] = t.hours
            st['ptab_f'].array[...] = t.ptab_f
            st['ptab_d'].array[...] = t.ptab_d
            st['ptab_s'].array[...] = t.ptab_s
            st['ray_off'].array[:n] = t.ray_off
            st['ray_dz'].array[:n] = t.ray_dz
            for key in ('hours', 'ptab_f', 'ptab_d', 'ptab_s'):
                rt.call('sw_memcpy_h2d', b[key].ptr, st[key].ptr, st[key].nbytes, h)
            rt.call('sw_memcpy_h2d', b['ray_off'].ptr, st['ray_off'].ptr, n * 8, h)
            rt.call('sw_memcpy_h2d', b['ray_dz'].ptr, st['ray_dz'].ptr, n * 8, h)
        self._stage_event[slot].record(self.stream)
        self._stage_used[slot] = True

    def reset_state(self):
        for d in self.domains:
            d.dev['lag'].zero(self.stream)

    # ------------------------------------------------------------------ launches
    def _launch_hours(self, h0, h1):
        rt.call('sw_day_forward' if self.resident else 'sw_hours_forward', self.dom_dev.ptr, self.block_start.ptr,
                len(self.domains), self.total_blocks, h0, h1, self.gvf_off.ptr, self.gvf_bounds.ptr, self.first,
                self.second, self.stream.handle)

    def run_hours(self, h0, h1):
        """Enqueue hours [h0, h1) of the loaded day (the lag state carries over from hour h0 - 1)."""
        if not 0 <= h0 <= h1 <= self.n_hours:
            raise ValueError(f'hours [{h0}, {h1}) outside the day of {self.n_hours}')
        self._launch_hours(int(h0), int(h1))

    def forward_day(self, use_graph=False):
        """Enqueue all hours of the loaded day (state must be reset by the caller per day)."""
        if use_graph:
            if self.graph is None:
                self.graph = rt.Graph.capture(self.stream, lambda: self._launch_hours(0, self.n_hours))
            self.graph.launch(self.stream)
        else:
            self._launch_hours(0, self.n_hours)

    def hour(self, h, full=False):
        """Both passes of one hour (for step-wise use and tests)."""
        rt.call('sw_hour_shadow', self.dom_dev.ptr, self.block_start.ptr, len(self.domains),
                self.total_blocks, h, int(full), self.stream.handle)
        rt.call('sw_hour_pixels', self.dom_dev.ptr, self.block_start.ptr, len(self.domains),
                self.total_blocks, h, self.gvf_off.ptr, self.gvf_bounds.ptr, self.first, self.second,
                int(full), self.stream.handle)

    def hour_masks(self, k=0):
        """Host mask bytes [rows, cols] of domain k from the last single-hour launch (``hour``):
        bit 0 not building-shadowed, 1 no canopy shadow, 2 vbshvegsh, 3 sunlit wall, 4 wall facing
        away from the sun."""
        self.stream.sync()
        d = self.domains[k]
        return self.buf[k]['masks'].to_host()[0].reshape(d.shape)

    def outputs(self, k=0):
        """Host (Tmrt, UTCI) [n_hours, rows, cols] of domain k for the last day run."""
        self.stream.sync()
        b, d = self.buf[k], self.domains[k]
        return (b['tmrt'].to_host().reshape(self.n_hours, *d.shape),
                b['utci'].to_host().reshape(self.n_hours, *d.shape))

    def extra_outputs(self, k=0):
        """Host {name: [n_hours, rows, cols]} of the extra planes (engine built with extra_outputs)."""
        self.stream.sync()
        b, d = self.buf[k], self.domains[k]
        e = b['extra'].to_host().reshape(self.n_hours, rt.SW_EXTRA_PLANES, *d.shape)
        return {name: e[:, i] for i, name in enumerate(rt.EXTRA_NAMES)}


class _Pending:
    def __init__(self, futures):
        self.futures = futures

    def result(self):
        return [f.result() for f in self.futures]


class _Lazy:
    def __init__(self, build):
        self.build = build

    def result(self):
        return self.build()


def _warm(_):
    return os.getpid()


def _day_tables(rows, sun, location, utc, latitude, shape, scale, amaxvalue, landcover, wall_params, constants,
                psi):
    """One domain's day tables (worker-process entry point as well)."""
    s = sun if sun is not None else met_sun(rows, location, utc)
    return build_day(rows, s, latitude, shape, scale, amaxvalue, landcover=landcover, wall_params=wall_params,
                     constants=constants, psi=psi)


def run_day(domain, met=None, sun=None, use_graph=False):
    """One day of Tmrt/UTCI for a synthetic.Domain (host arrays out)."""
    met = domain.met if met is None else met
    eng = Engine([domain], n_hours=met.shape[0])
    t = eng.tables(met=[met], sun=None if sun is None else [sun])
    eng.load(t)
    eng.reset_state()
    eng.forward_day(use_graph)
    return eng.outputs(0)
