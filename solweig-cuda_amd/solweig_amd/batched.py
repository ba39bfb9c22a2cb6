"""Batched day forward: every hour of a day, for one or many uneven domains, on one stream.

The solar position of every forcing row comes from the device (FP64 SPA, sw_sun_position, once
per domain for its whole forcing). Per day the host builds each domain's hour tables
(sky.build_day) and uploads them into buffers of fixed capacity; the device then runs the day
either resident (sw_day_forward: the shadow pass of every hour in one launch, then one pixel pass
in which each tile runs through the hours with its lag state in registers) or as one shadow and
one pixel launch per hour (sw_hours_forward), each launch covering the tiles of all domains
together. Lag state is cleared at each day start (series.py:425-436). Because buffer addresses
never change, a day's launches can be captured once as a hipGraph and replayed for every further
day after the tables are refreshed. Consecutive days can run pipelined (run_days, day_start /
day_next): each day's all-hours shadow pass on a side stream beside the previous day's pixel pass,
the days alternating between two table and record sets.
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
    def __init__(self, domains, n_hours=24, stream=None, keep_outputs=True, extra_outputs=False, resident=True,
                 pipelined=False):
        """``extra_outputs``: also keep each hour's Kdown, Kup, Ldown, Lup and Shadow planes
        ([n_hours][5][P] per domain, rt.EXTRA_NAMES order), the optional series outputs.
        ``resident``: run a day as two launches (sw_day_forward: the shadow pass of every hour into
        per-hour record and mask planes, 9 B per pixel-hour, then one pixel pass in which each tile
        runs through the hours with its lag state in registers) instead of two launches per hour
        (sw_hours_forward); an engine built with ``resident=False`` keeps one record slot and
        cannot switch to the resident forms.
        ``pipelined``: multi-day callers (series, tiling, bench) run consecutive days through
        ``run_days``, each day's shadow pass alongside the previous day's pixel pass (resident
        form only); ``forward_day`` always runs one day on its own."""
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
                     diff_tab=rt.DeviceArray((rt.diff_table_bytes(self.n_hours),), np.uint8),
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
            # the day's diffuse sums of every pixel-hour (sw_day_diffuse), read by the pixel passes
            b['diffuse'] = rt.DeviceArray((self.n_hours, d.P, 2), np.float32)
            self.buf.append(b)
        self.resident = bool(resident)
        self.pipelined = bool(pipelined) and self.resident
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
                             diff_tab=rt.PinnedArray(b['diff_tab'].shape, np.uint8),
                             ray_off=rt.PinnedArray((b['cap'], 2), np.int32),
                             ray_dz=rt.PinnedArray((b['cap'],), np.float64)) for b in self.buf]
                       for _ in range(2)]
        self._stage_event = [rt.Event(), rt.Event()]
        self._stage_used = [False, False]
        self._slot = 0
        self.graph = None
        # pipelined days (day_start / day_next): a second table and record set, a side stream for
        # the shadow passes, and the set the current day's tables and records are in
        self._set1 = None
        self.dom_devs = [self.dom_dev]
        self._side = None
        self._pipe_set = None
        self._pipe_graphs = {}
        self.groups_beside_shadow = 3
        # per table set: the diffuse sums (sw_day_diffuse) of the loaded day are enqueued
        self._diffuse_ok = [False, False]

    def _write_domains(self):
        """Upload the domain records and the tile prefix (after a change of tile widths)."""
        starts = np.zeros(len(self.domains) + 1, dtype=np.int32)
        starts[1:] = np.cumsum([d.tiles for d in self.domains])
        self.total_blocks = int(starts[-1])
        self.block_start.upload(starts, self.stream)
        sets = [(self.buf, self.dom_dev)]
        if getattr(self, '_set1', None) is not None:
            sets.append((self._set1, self.dom_devs[1]))
        for bufs, dev in sets:
            for i, (d, b, o) in enumerate(zip(self.domains, bufs, self.buf)):
                self.structs[i] = d.struct(hours=b['hours'].ptr, ptab_f=b['ptab_f'].ptr, ptab_d=b['ptab_d'].ptr,
                                           ptab_s=b['ptab_s'].ptr, ray_off=b['ray_off'].ptr, ray_dz=b['ray_dz'].ptr,
                                           diffuse_tab=b['diff_tab'].ptr, diffuse=b['diffuse'].ptr,
                                           tmrt=o['tmrt'].ptr if 'tmrt' in o else 0,
                                           utci=o['utci'].ptr if 'utci' in o else 0,
                                           full=o['extra'].ptr if 'extra' in o else 0,
                                           rec=b['rec'].ptr, masks=b['masks'].ptr)
            rt.call('sw_memcpy_h2d', dev.ptr, ctypes.addressof(self.structs), ctypes.sizeof(self.structs),
                    self.stream.handle)
            self.stream.sync()
        self.graph = None                                # launch geometry changed
        self._pipe_graphs = {}

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

    def load(self, tables, set_index=0):
        """Queue the upload of one day's tables on the engine's stream (fixed device addresses, so
        a captured graph stays valid). The copies run after every launch queued before them, so
        the host may build and load the next day while the device still runs this one.
        ``set_index`` 1: into the second table set of the pipelined days (day_start / day_next).
        Refused while a run of pipelined days is open: its days own both sets."""
        self._refuse_open()
        self._load(tables, set_index)

    def _load(self, tables, set_index=0):
        bufs = self.buf if set_index == 0 else self._set1
        for t, b in zip(tables, bufs):
            if t.n_hours != self.n_hours:
                raise ValueError(f'expected {self.n_hours} hours, got {t.n_hours}')
            if t.ray_dz.shape[0] > b['cap']:
                raise ValueError('ray tables exceed their buffer')
        self._diffuse_ok[set_index] = False
        slot = self._slot
        self._slot ^= 1
        if self._stage_used[slot]:
            self._stage_event[slot].sync()        # the copies out of this slot have finished
        h = self.stream.handle
        for t, b, st in zip(tables, bufs, self._stage[slot]):
            n = t.ray_dz.shape[0]
            st['hours'].array[...] = t.hours
            st['ptab_f'].array[...] = t.ptab_f
            st['ptab_d'].array[...] = t.ptab_d
            st['ptab_s'].array[...] = t.ptab_s
            st['diff_tab'].array[...] = t.diffuse
            st['ray_off'].array[:n] = t.ray_off
            st['ray_dz'].array[:n] = t.ray_dz
            for key in ('hours', 'ptab_f', 'ptab_d', 'ptab_s', 'diff_tab'):
                rt.call('sw_memcpy_h2d', b[key].ptr, st[key].ptr, st[key].nbytes, h)
            rt.call('sw_memcpy_h2d', b['ray_off'].ptr, st['ray_off'].ptr, n * 8, h)
            rt.call('sw_memcpy_h2d', b['ray_dz'].ptr, st['ray_dz'].ptr, n * 8, h)
        self._stage_event[slot].record(self.stream)
        self._stage_used[slot] = True

    def reset_state(self):
        for d in self.domains:
            d.dev['lag'].zero(self.stream)

    # ------------------------------------------------------------------ launches
    def _check_resident(self):
        if self.buf[0]['rec'].shape[0] != self.n_hours:
            raise RuntimeError('this engine was built with resident=False: its record planes hold one hour')

    def _refuse_open(self):
        if self.day_open:
            raise RuntimeError('a run of pipelined days is open (day_start without its last day_next)')

    def diffuse(self, k=0, stream=None):
        """Enqueue the diffuse sums of every pixel-hour of the day loaded in table set k
        (sw_day_diffuse), which the pixel passes of that day read; ``stream``: the engine's."""
        rt.call('sw_day_diffuse', self.dom_devs[k].ptr, self.block_start.ptr, len(self.domains), self.total_blocks,
                self.n_hours, (stream or self.stream).handle)
        self._diffuse_ok[k] = True

    def _launch_hours(self, h0, h1):
        self._refuse_open()
        if self.resident:
            self._check_resident()
        self.diffuse()
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

    # ------------------------------------------------------------------ pipelined days
    # Shadows depend only on the static rasters and the day's tables, never on the lag state, so
    # the all-hours shadow pass of day k + 1 can run while the pixel pass of day k runs: the two
    # kernels are bound differently (the march by scalar window tests and gather latency, the
    # pixel pass by vector issue) and fill each other's idle issue slots. Days alternate between
    # two table and record sets; the shadow passes run on a side stream.
    def _ensure_pipeline(self):
        self._check_resident()
        if self._set1 is not None:
            return
        self._set1 = []
        for d, b in zip(self.domains, self.buf):
            self._set1.append(dict(hours=rt.DeviceArray(b['hours'].shape, np.float64),
                                   ptab_f=rt.DeviceArray(b['ptab_f'].shape, np.float32),
                                   ptab_d=rt.DeviceArray(b['ptab_d'].shape, np.float64),
                                   ptab_s=rt.DeviceArray(b['ptab_s'].shape, np.uint32),
                                   diff_tab=rt.DeviceArray(b['diff_tab'].shape, np.uint8),
                                   ray_off=rt.DeviceArray((b['cap'], 2), np.int32),
                                   ray_dz=rt.DeviceArray((b['cap'],), np.float64), cap=b['cap'],
                                   rec=rt.DeviceArray((self.n_hours, d.P, 2), np.float32),
                                   masks=rt.DeviceArray((self.n_hours, d.P), np.uint8),
                                   diffuse=rt.DeviceArray((self.n_hours, d.P, 2), np.float32)))
        self.dom_devs = [self.dom_dev, rt.DeviceArray((ctypes.sizeof(self.structs),), np.uint8)]
        self._side = rt.Stream()
        self._fork, self._join = rt.Event(), rt.Event()
        self._write_domains()

    def _args(self, k):
        return (self.dom_devs[k].ptr, self.block_start.ptr, len(self.domains), self.total_blocks, 0, self.n_hours)

    def _shadow_pass(self, k, stream):
        """The all-hours shadow pass of set k and the diffuse sums of its day."""
        rt.call('sw_day_shadow', *self._args(k), stream.handle)
        self.diffuse(k, stream)

    def _pixel_pass(self, k, stream, beside_shadow=False):
        """Pixel pass of set k; ``beside_shadow``: the next day's shadow pass and diffuse sums run
        on the side stream meanwhile, so at most ``groups_beside_shadow`` pixel workgroups share a
        CU (at 2000^2: 7.31-7.36 ms per pipelined day with 3 per CU, 7.37-7.44 with 4, 7.49-7.50
        with 5 or no cap, 8.9 with 2). Splitting the day's pixel pass into a capped part beside
        the shadow pass and an uncapped rest gave no gain (7.32-7.40 ms for splits at hours 8-18)."""
        if beside_shadow and self.groups_beside_shadow:
            rt.call('sw_day_pixels_capped', *self._args(k), self.gvf_off.ptr, self.gvf_bounds.ptr, self.first,
                    self.second, self.groups_beside_shadow, stream.handle)
        else:
            rt.call('sw_day_pixels', *self._args(k), self.gvf_off.ptr, self.gvf_bounds.ptr, self.first,
                    self.second, stream.handle)

    def _overlapped(self, k, shadow_next):
        """The pixel pass of set k on the engine's stream and, with ``shadow_next``, the shadow pass
        of set 1 - k on the side stream, joined back into the engine's stream."""
        if shadow_next:
            self._fork.record(self.stream)
            self._side.wait(self._fork)
            self._shadow_pass(1 - k, self._side)
            self._join.record(self._side)
        self._pixel_pass(k, self.stream, beside_shadow=shadow_next)
        if shadow_next:
            self.stream.wait(self._join)

    def _capture_pipeline(self):
        """Capture the four graphs of day_next (set 0 or 1, with or without the next shadow pass)."""
        self._ensure_pipeline()
        for k in (0, 1):
            for more in (True, False):
                if (k, more) not in self._pipe_graphs:
                    self._pipe_graphs[(k, more)] = rt.Graph.capture(
                        self.stream, lambda k=k, more=more: self._overlapped(k, more))

    @property
    def day_open(self):
        """True between day_start and the day_next that had no next day."""
        return self._pipe_set is not None

    def day_start(self, tables):
        """First day of a run of pipelined days (resident launch form): load its tables and enqueue
        its shadow pass. Then call ``day_next`` once per day. While a run is open (``day_open``) the
        engine's table sets belong to it: ``load``, ``forward_day`` and ``run_hours`` would
        overwrite the tables or records of a day still in flight."""
        self._ensure_pipeline()
        self._pipe_set = 0
        self._load(tables, 0)
        self._shadow_pass(0, self.stream)

    def day_next(self, next_tables=None, use_graph=False):
        """Enqueue the pixel pass of the current day (lag state cleared first) on the engine's
        stream; with ``next_tables`` the next day's tables go into the other set and its shadow pass
        runs alongside, on the side stream. When this returns, the current day's outputs are
        complete in the engine's stream order, until the next day_next overwrites them. With
        ``use_graph`` the two passes are replayed from a captured graph (one per set and form)."""
        if self._pipe_set is None:
            raise RuntimeError('day_next without day_start')
        k, more = self._pipe_set, next_tables is not None
        if more:
            self._load(next_tables, 1 - k)
        self.reset_state()
        if use_graph:
            self._capture_pipeline()
            self._pipe_graphs[(k, more)].launch(self.stream)
        else:
            self._overlapped(k, more)
        self._pipe_set = 1 - k if more else None

    def run_days(self, tables, use_graph=False, after_day=None):
        """Run consecutive days through day_start / day_next. ``tables``: an iterable of day tables
        (or of objects with ``result()``, e.g. ``tables_async``), consumed one day ahead;
        ``after_day(i)`` is called when day i's pixel pass is enqueued (its outputs are then
        complete in the engine's stream order). Returns the number of days run."""
        it = iter(tables)

        def take():
            t = next(it, None)
            return t.result() if hasattr(t, 'result') else t

        if use_graph:
            self._capture_pipeline()
        cur, n = take(), 0
        if cur is None:
            return 0
        self.day_start(cur)
        while True:
            nxt = take()
            self.day_next(nxt, use_graph=use_graph)
            if after_day is not None:
                after_day(n)
            n += 1
            if nxt is None:
                return n

    def hour(self, h, full=False):
        """Both passes of one hour (for step-wise use and tests)."""
        self._refuse_open()
        if not full and not self._diffuse_ok[0]:
            self.diffuse()                               # once per loaded day
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
