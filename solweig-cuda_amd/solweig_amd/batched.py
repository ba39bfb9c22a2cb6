"""Batched day forward: every hour of a day, for one or many uneven domains, on one stream.

Per day the host builds each domain's hour tables (sky.build_day), uploads them into buffers of
fixed capacity, and the device runs, per hour, one shadow launch and one pixel launch that cover
the tiles of all domains together (sw_hours_forward). Lag state stays on the device across the
hours of a day and is cleared at each day start (series.py:425-436). Because buffer addresses
never change, the 2 x n_hours launches of a day can be captured once as a hipGraph and replayed
for every further day after the tables are refreshed.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import runtime as rt
from .sky import CONSTANTS, STEP_BLOCK, build_day, gvf_geometry
from .static import PreparedDomain
from .sun import met_sun


class Engine:
    def __init__(self, domains, n_hours=24, stream=None, keep_outputs=True):
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
            self.buf.append(b)
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
                                       utci=b['utci'].ptr if 'utci' in b else 0)
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
    def tables(self, day_index=0, met=None, sun=None, constants=None, psi=None):
        """Host tables of one day for every domain. ``met`` / ``sun`` override per domain;
        ``constants`` (sky.CONSTANTS-like) and ``psi`` (canopy transmissivity of the hour's
        shadow, default by season) override the model constants."""
        out = []
        for k, d in enumerate(self.domains):
            dm = d.domain
            rows = met[k] if met is not None else dm.met[day_index * self.n_hours:(day_index + 1) * self.n_hours]
            s = sun[k] if sun is not None else met_sun(rows, dm.location, dm.utc)
            out.append(build_day(rows, s, dm.latitude, d.shape, dm.scale, d.amaxvalue,
                                 landcover=d.landcover, wall_params=d.wall_params,
                                 constants=constants or CONSTANTS, psi=psi))
        return out

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
        rt.call('sw_hours_forward', self.dom_dev.ptr, self.block_start.ptr, len(self.domains),
                self.total_blocks, h0, h1, self.gvf_off.ptr, self.gvf_bounds.ptr, self.first,
                self.second, self.stream.handle)

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

    def outputs(self, k=0):
        """Host (Tmrt, UTCI) [n_hours, rows, cols] of domain k for the last day run."""
        self.stream.sync()
        b, d = self.buf[k], self.domains[k]
        return (b['tmrt'].to_host().reshape(self.n_hours, *d.shape),
                b['utci'].to_host().reshape(self.n_hours, *d.shape))


def run_day(domain, met=None, sun=None, use_graph=False):
    """One day of Tmrt/UTCI for a synthetic.Domain (host arrays out)."""
    met = domain.met if met is None else met
    eng = Engine([domain], n_hours=met.shape[0])
    t = eng.tables(met=[met], sun=None if sun is None else [sun])
    eng.load(t)
    eng.reset_state()
    eng.forward_day(use_graph)
    return eng.outputs(0)
