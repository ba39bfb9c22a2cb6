"""Benchmark: Tmrt+UTCI pixel-hours/s of the batched day forward on one MI355X.

A step is one day (24 hourly steps) of the batched forward over the 2000x2000 c2 workload
(seeded synthetic morphology, its 30-day forcing cycled; the default 30 steps time the whole
720-hour month, each day once): per step the host builds the day's
tables (clearness, Perez, patch rows, rays; the solar position comes from the device) in CPU
worker processes, the engine uploads them, and the device replays the day's launches as captured
hipGraphs, in the launch form that a timed comparison before the warm-up picks:
* pipelined: resident days whose all-hours shadow pass runs on a side stream beside the previous
  day's pixel pass (batched.Engine.run_days; every timed run starts with a lone shadow pass and
  ends with a lone pixel pass, so all of its days' work lies inside the timed region);
* resident: the shadow pass of all 24 hours, then the pixel pass in which every tile runs through
  the 24 hours with its lag state in registers, one day after the other;
* hourly: two launches per hour.
Tmrt and UTCI stay in device memory. Static preparation (walls, sky view) runs before the timed
region and is reported apart.

Beside the headline the JSON line carries:
* ``config.eager``: the same days launched eagerly with HIP events around both launches, which give
  the per-kernel durations of ``roofline`` (dominant kernel: the pixel pass);
* ``config.launch_form`` / ``other_launch_form``: the timed choice between the three forms, and the
  same days in the forms not chosen;
* ``config.plain_forward``: the per-step drop-in API (host tables, uploads, two full-output
  launches, the UTCI kernel and a stream sync every hour), and under ``stepwise`` the per-step
  forward in the reference's own decomposition (about 20 standalone sub-model launches, each
  followed by a sync, per daytime hour) with the batched / stepwise ratio;
* ``config.forward_adjoint``: forward vs forward + adjoint on the 1000x1000 differentiation workload;
* ``config.c4``: the 64-domain uneven batch, pipelined resident days (one launch per pass and day
  over all domains), tables built inside its timed region;
* ``cpu_baseline``: the FP64 NumPy oracle (a port of the UMEP step) on the 200x200 c0 workload,
  its full 24 h under both land-cover readings, on one host core.

``--gpus N`` with N > 1 is refused: the benchmark runs on one GPU.
"""
from __future__ import annotations

import argparse
from collections import deque
import datetime
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'solweig-cuda_amd'))

import numpy as np  # noqa: E402

# Algorithmic bytes per pixel-hour of the forward: the static inputs at their minimal exact size,
# read once, plus the outputs. Shadow pass: DSM, canopy, walls, aspect (16 B). Pixel pass: ground
# bit (0.125), 15 SVF planes FP32 (60), tan(asvf) and tan(svfalfa) FP64 (16), 3 x 153 visibility
# bits (57.375), Tmrt + UTCI FP32 (8): 141.5 B. (The diffuse sums of sw_day_diffuse, like the shadow
# pass's records, are intermediates: the pre-pass reads the sky and vegetation bits, the pixel pass
# the building bits.)
B_SHADOW = 16.0
B_PIXELS = 141.5
B_ALG = B_SHADOW + B_PIXELS
HBM_PEAK = 8.0e12      # MI355X HBM3E bytes/s
DIFF_EVENT = 3 * 24    # per eager day: the events around the diffuse-sum launch (see run_days)
# HBM bytes per pixel-pass launch at 2000x2000: rocprofv3 FETCH_SIZE scaled by the k_copy
# calibration + WRITE_SIZE, from the profile named in TRAFFIC_SOURCE (profile values, not measured by
# this run): k_day_pixels (one launch for 24 hours), k_hour_pixels (mean over a day's 24 launches).
TRAFFIC_2000 = {'k_day_pixels': 23.98e9, 'k_hour_pixels': 0.796e9}
TRAFFIC_SOURCE = 'profiles/k_day_pixels_c2_2000x2000.md'
# VALU instructions (wave64) per pixel-pass launch at 2000x2000, SQ_INSTS_VALU x SQ_WAVES from the same
# profile, and the single-issue VALU rate of MI355X (256 CUs x 2.4 GHz, one wave64 instruction per CU
# per cycle without the packed FP32 forms): the ceiling the pixel pass actually runs against.
VALU_2000 = {'k_day_pixels': 45554 * 62500, 'k_hour_pixels': 1873 * 62500}
VALU_PEAK = 256 * 2.4e9
FORM_TEXT = {
    'pipelined': 'resident days pipelined over two streams (each day\'s all-hours shadow pass and its diffuse '
                 'sums on the matrix cores beside the previous day\'s pixel pass, batched.Engine.run_days; the '
                 'timed run starts with a lone shadow pass and ends with a lone pixel pass)',
    'resident': 'the resident day (sw_day_diffuse, then sw_day_forward)',
    'hourly': 'the diffuse sums of the day, then two launches per hour (sw_hours_forward)',
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    model = platform.processor() or ''
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    aff = sorted(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else []
    return model, os.cpu_count(), aff


def cpu_baseline():
    """The FP64 NumPy oracle on the 200x200 c0 workload: its full 24 h under both readings of the
    land cover (all paved, and the default maps), after each reading's static sky-view preparation
    (timed apart), pinned to one host core."""
    model, count, aff = host_info()
    core = aff[0] if aff else 0
    try:
        os.sched_setaffinity(0, {core})
    except (AttributeError, OSError):
        pass
    from solweig_amd import synthetic, sun as psun
    from oracle import day as oday
    runs, total_px_h, total_s = [], 0, 0.0
    for paved in (True, False):
        d = synthetic.make_c0(landcover=paved)
        s = psun.met_sun(d.met, d.location, d.utc)
        t0 = time.perf_counter()
        p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude, landcover=d.landcover)
        t_static = time.perf_counter() - t0
        t0 = time.perf_counter()
        oday.run_day(p, d.met, s)
        dt = time.perf_counter() - t0
        px_h = d.dsm.size * d.met.shape[0]
        total_px_h += px_h
        total_s += dt
        runs.append(dict(land_cover='all paved (code 1)' if paved else 'default maps', pixel_hours=px_h,
                         seconds=dt, static_svf_seconds=t_static))
        log(f'[bench] cpu baseline c0 {"paved" if paved else "default"}: {px_h} px-h in {dt:.1f} s '
            f'(static sky view {t_static:.1f} s)')
    return dict(value=total_px_h / total_s, unit='pixel-hours/s', cores=1, kind='port',
                sample=f'oracle FP64 NumPy UMEP step on c0 200x200, all 24 h, both land-cover readings '
                       f'({total_px_h} px-h in {total_s:.1f} s; static sky-view preparation excluded)',
                cpu_model=model, os_cpu_count=count, affinity=[core], affinity_before=aff, runs=runs)


def static_leg(dom):
    """Static preparation on the device, timed with events: wall heights and aspect from the DSM
    (the reference's calculate_dsm_walls; the domain's walls and aspect are replaced by these, as
    run_timeseries does), the sky-view march and reduction, the static sky sums."""
    from solweig_amd import runtime as rt, static as pstatic, svf as psvf, walls as pwalls
    stream = rt.Stream()
    ev = [rt.Event() for _ in range(4)]
    dsm64 = rt.DeviceArray.from_host(dom.dsm.astype(np.float64))
    pwalls.dsm_walls_device(dsm64, 1.0, 3.0, stream=stream)          # warm-up (tables, code load)
    ev[0].record(stream)
    w, a = pwalls.dsm_walls_device(dsm64, 1.0, 3.0, stream=stream)
    ev[1].record(stream)
    stream.sync()
    out = dict(walls_ms=ev[0].elapsed_ms(ev[1]))
    dom.walls, dom.aspect = w.to_host().astype(np.float32), a.to_host().astype(np.float32)
    out['wall_pixels'] = int((dom.walls > 0).sum())
    prep = pstatic.PreparedDomain(dom, stream=stream)                 # includes one sky view
    d = prep.dev
    sv = psvf.sky_view(d['dsm'], d['trees'], prep.dsm_raise, prep.shape, prep.amaxvalue, dom.scale, prep.trans,
                       stream, events=ev[:3], hmax=d['hmax'], march_trees=d['march_trees'])
    out['svf_visibility_ms'], out['svf_reduce_ms'] = ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])
    ev[0].record(stream)
    pstatic.sky_sums(sv.bits, prep.P, stream)
    ev[1].record(stream)
    stream.sync()
    out['sky_sums_ms'] = ev[0].elapsed_ms(ev[1])
    del prep, sv
    return out


def adjoint_leg(size=1000, repeats=3):
    """Forward vs forward + adjoint on the 1000x1000 differentiation workload (C3 morphology with
    per-pixel albedo U(0.1, 0.4) and emissivity U(0.90, 0.98) maps; J = mean ground UTCI at 14:00)."""
    from solweig_amd import adjoint, synthetic
    d = synthetic.make_c3(size=size)
    prob = adjoint.UtciGradient(d, objective_hour=14)
    eng = prob.eng

    # every timed call starts from a new parameter set (the hour tables are rebuilt); the gradient
    # after an objective of the same parameters reuses that objective's tables
    def tables():
        prob.clear_cache()
        return prob._tables()

    def forward_day():
        _, tabs = tables()
        eng.load(tabs)
        eng.reset_state()
        eng.forward_day()
        eng.stream.sync()

    def forward_adjoint():
        prob.clear_cache()
        t0 = time.perf_counter()
        prob.objective()
        t1 = time.perf_counter()
        g = prob.gradient()
        return (t1 - t0) * 1e3, (time.perf_counter() - t1) * 1e3, g

    def timed(fn):
        fn()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3, out

    t_tab, _ = timed(tables)
    t_fwd, _ = timed(forward_day)
    forward_adjoint()
    runs = [forward_adjoint() for _ in range(repeats)]
    t_obj = float(np.median([r[0] for r in runs]))
    t_grad = float(np.median([r[1] for r in runs]))
    g = runs[-1][2]
    log(f'[bench] adjoint leg: hour tables {t_tab:.1f} ms, forward day {t_fwd:.1f} ms, objective {t_obj:.1f} ms, '
        f'gradient {t_grad:.1f} ms')
    return dict(workload=f'c3 {size}x{size}, 24 h forcing, J = mean UTCI over ground pixels at 14:00; '
                         f'gradient w.r.t. albedo and emissivity maps, wall albedo, transmissivity and hourly '
                         f'Ta, RH, wind, radG',
                tables_ms=t_tab, forward_day_ms=t_fwd, objective_ms=t_obj, gradient_ms=t_grad,
                gradient_over_forward=t_grad / t_fwd, forward_plus_adjoint_over_forward=(t_obj + t_grad) / t_fwd,
                without_tables=(t_obj + t_grad - t_tab) / (t_fwd - t_tab),
                J=g['J'], hours=list(g['hours']))


class DayPipeline:
    """Day tables from the table workers, kept ``depth`` days ahead of the device over one
    continuous sequence of days (warm-up, timed and event-timed days alike), so that the timed
    days see the steady state: while the device runs a day, the workers build the following ones."""

    def __init__(self, eng, day_indices):
        self.eng, self.days, self.nxt = eng, list(day_indices), 0
        self.depth = max(2, getattr(eng, 'workers', 0))
        self.pending = deque()
        self._fill()

    def _fill(self):
        while self.nxt < len(self.days) and len(self.pending) < self.depth:
            self.pending.append(self.eng.tables_async(day_index=self.days[self.nxt]))
            self.nxt += 1

    def next(self):
        tabs = self.pending.popleft().result()
        self._fill()
        return tabs


def run_days(eng, pipe, n, use_graph, events=None):
    """``n`` days from the pipeline: per day the upload (stream-ordered behind the previous day),
    the lag reset and the 24 hours. With ``events`` (one list of 3 x 24 + 1 per day) every launch
    is bracketed by HIP events: the two launches of a resident engine's day (events 0-2), or the
    shadow and pixel launch of every hour (events 3h .. 3h + 2)."""
    from solweig_amd import runtime as rt
    used = []
    if events is None and eng.pipelined:
        # consecutive days with each day's shadow pass beside the previous day's pixel pass; the
        # run starts with the first day's shadow pass alone and ends with the last pixel pass alone
        def days():
            for _ in range(n):
                used.append(pipe.next())
                yield used[-1]
        eng.run_days(days(), use_graph=use_graph)
        eng.stream.sync()
        return used
    for k in range(n):
        tabs = pipe.next()
        eng.load(tabs)
        eng.reset_state()
        if events is None:
            eng.forward_day(use_graph=use_graph)
        elif eng.resident:
            ev = events[k]
            args = (eng.dom_dev.ptr, eng.block_start.ptr, len(eng.domains), eng.total_blocks, 0, eng.n_hours)
            # shadow pass (events 0 -> 1), diffuse sums (1 -> DIFF_EVENT), pixel pass (DIFF_EVENT -> 2)
            ev[0].record(eng.stream)
            rt.call('sw_day_shadow', *args, eng.stream.handle)
            ev[1].record(eng.stream)
            eng.diffuse()
            ev[DIFF_EVENT].record(eng.stream)
            rt.call('sw_day_pixels', *args, eng.gvf_off.ptr, eng.gvf_bounds.ptr, eng.first, eng.second,
                    eng.stream.handle)
            ev[2].record(eng.stream)
        else:
            ev = events[k]
            # the diffuse sums of the day (events DIFF_EVENT -> DIFF_EVENT + 1), then per hour the
            # shadow pass (3h -> 3h + 1) and the pixel pass (3h + 1 -> 3h + 2)
            ev[DIFF_EVENT].record(eng.stream)
            eng.diffuse()
            ev[DIFF_EVENT + 1].record(eng.stream)
            for h in range(eng.n_hours):
                ev[3 * h].record(eng.stream)
                rt.call('sw_hour_shadow', eng.dom_dev.ptr, eng.block_start.ptr, len(eng.domains), eng.total_blocks,
                        h, 0, eng.stream.handle)
                ev[3 * h + 1].record(eng.stream)
                rt.call('sw_hour_pixels', eng.dom_dev.ptr, eng.block_start.ptr, len(eng.domains),
                        eng.total_blocks, h, eng.gvf_off.ptr, eng.gvf_bounds.ptr, eng.first, eng.second, 0,
                        eng.stream.handle)
                ev[3 * h + 2].record(eng.stream)
        used.append(tabs)
    eng.stream.sync()
    return used


def c4_leg(args, workers):
    """The uneven-size batch: 64 domains (sides U_int[256, 1536], latitudes U(0, 60) N, each with
    its own morphology and forcing) in one engine: pipelined resident days, one launch per pass and
    day over all of them (each day's shadow pass beside the previous day's pixel pass). A step is
    one day; each day's tables (64 domains) are built by the table workers inside the
    timed region. The smallest and the largest domain are run again alone and must match their part
    of the batch bit for bit."""
    from solweig_amd import runtime as rt, synthetic, batched
    t0 = time.perf_counter()
    doms = synthetic.make_c4(count=args.c4_count)
    sides = [d.shape[0] for d in doms]
    P = sum(d.dsm.size for d in doms)
    t_gen = time.perf_counter() - t0
    log(f'[bench] c4: {len(doms)} domains, sides {min(sides)}..{max(sides)}, {P / 1e6:.1f} Mpx '
        f'(generated in {t_gen:.1f} s)')
    t0 = time.perf_counter()
    eng = batched.Engine(doms, n_hours=24, pipelined=True)
    rt.synchronize()
    t_prep = time.perf_counter() - t0
    eng.set_tiles(args.tile or 16)
    eng.workers = 0 if args.no_workers else eng.start_table_workers(workers)
    log(f'[bench] c4 static preparation {t_prep:.1f} s, {eng.total_blocks} tiles, {eng.workers} table workers')
    t0 = time.perf_counter()
    tabs = eng.tables()
    t_tab_serial = time.perf_counter() - t0
    warm = max(1, args.warmup)
    pipe = DayPipeline(eng, [0] * (warm + args.c4_steps))
    use_graph = not args.eager
    run_days(eng, pipe, warm, use_graph=use_graph)
    rt.synchronize()
    t0 = time.perf_counter()
    run_days(eng, pipe, args.c4_steps, use_graph=use_graph)
    rt.synchronize()
    dt = time.perf_counter() - t0
    value = P * 24 * args.c4_steps / dt
    log(f'[bench] c4: {dt / args.c4_steps * 1e3:.1f} ms/day, {value:.3e} px-h/s (tables built in the timed region; '
        f'one core would take {t_tab_serial * 1e3:.0f} ms/day)')
    checks = []
    for k in (int(np.argmin(sides)), int(np.argmax(sides))):
        tm, ut = eng.outputs(k)
        alone = batched.Engine([doms[k]], n_hours=24)
        alone.set_tiles(args.tile or 16)
        alone.load([tabs[k]])
        alone.reset_state()
        alone.forward_day()
        tm1, ut1 = alone.outputs(0)
        same = bool(np.array_equal(tm, tm1, equal_nan=True) and np.array_equal(ut, ut1, equal_nan=True))
        checks.append(dict(domain=k, side=sides[k], identical_to_single_domain_run=same))
        log(f'[bench] c4 domain {k} ({sides[k]}^2) alone vs in the batch: identical {same}')
        del alone
    eng.close_table_workers()
    audit = c4_audit(eng, doms, sides, args.c4_audit)
    return dict(value=value, unit='pixel-hours/s', ms_per_day=dt / args.c4_steps * 1e3, steps=args.c4_steps,
                workload=f'c4 uneven batch: {len(doms)} domains, sides {min(sides)}..{max(sides)}, {P} pixels, '
                         f'24 h per step, pipelined resident days (one launch per pass and day over all '
                         f'domains{", replayed from hipGraphs" if use_graph else ""}), day tables built in the '
                         f'timed region',
                graph=use_graph,
                tile=args.tile or 16, generation_s=t_gen, static_prep_s=t_prep,
                host_tables_one_core_ms=t_tab_serial * 1e3, table_workers=eng.workers, checks=checks,
                fp64_audit=audit)


def c4_audit(eng, doms, sides, count):
    """FP32 batched outputs of c4 domains (smallest, median, largest, ...) against the FP64 step
    decomposition of the same day on the device (series.stepwise_day in FP64, which equals the FP64
    oracle to float32 rounding): max and RMS |dTmrt|, |dUTCI| in K over the pixel-hours."""
    from solweig_amd import series
    from solweig_amd.sun import met_sun
    order = list(np.argsort(sides, kind='stable'))
    picks = [order[0], order[len(order) // 2], order[-1]] + order[1:-1:max(1, len(order) // max(1, count - 3))]
    out = []
    for k in [int(i) for i in dict.fromkeys(picks)][:max(0, count)]:
        d = doms[k]
        tm, ut = eng.outputs(k)
        t0 = time.perf_counter()
        ctx = series.stepwise_context(d, dtype=np.float64)
        m = d.met[:24]
        t64, u64, _ = series.stepwise_day(*ctx, m, met_sun(m, d.location, d.utc))
        del ctx
        dt_ = np.abs(tm.astype(np.float64) - t64)
        g = np.isfinite(u64) & np.isfinite(ut)
        du = np.abs(ut[g].astype(np.float64) - u64[g])
        same_nan = bool(np.array_equal(np.isnan(ut), np.isnan(u64)))
        row = dict(domain=k, side=sides[k], pixel_hours=int(tm.size), tmrt_max=float(dt_.max()),
                   tmrt_rms=float(np.sqrt(np.mean(dt_ ** 2))), utci_max=float(du.max()),
                   utci_rms=float(np.sqrt(np.mean(du ** 2))), utci_nan_pattern_equal=same_nan,
                   seconds=time.perf_counter() - t0)
        out.append(row)
        log(f'[bench] c4 FP64 audit domain {k} ({sides[k]}^2): Tmrt max {row["tmrt_max"]:.2e} rms '
            f'{row["tmrt_rms"]:.2e} K, UTCI max {row["utci_max"]:.2e} rms {row["utci_rms"]:.2e} K, '
            f'NaN pattern equal {same_nan} ({row["seconds"]:.1f} s)')
    return out


def e2e_leg(repeats=3):
    """The complete workflow of series.run_timeseries on the 2000x2000 c2 rasters (.npy inputs):
    walls and aspect on the device, static preparation (sky view, sums, ground-view planes), the
    day loop and the outputs (UTCI and TMRT stacks written as .npy with their tag files under a
    temporary directory; the reference writes DEFLATE GeoTIFF). 1 day written and 30 days in the
    reference's no-write mode, each after one warm-up run as the median of ``repeats`` runs (the
    protocol of the reference's README.md:26-32); 30 days written once (23 GB of .npy). Context: the reference publishes 2.7 s (1 day) and 21 s
    (30 days) for this raster size on an RTX 4090 (README.md:17-24)."""
    import shutil
    import tempfile
    from pathlib import Path
    from solweig_amd import series, synthetic
    d = synthetic.make_c1(seed=2, size=2000, days=30, start=datetime.date(2025, 6, 6), name='c2')
    tmp = Path(tempfile.mkdtemp(prefix='solweig_e2e_', dir='/tmp' if os.path.isdir('/tmp') else None))
    try:
        for name in ('dsm', 'trees', 'dem'):
            np.save(tmp / f'{name}.npy', getattr(d, name))
        paths = [tmp / f'{name}.npy' for name in ('dsm', 'trees', 'dem')]

        def run(days, write, k):
            out = tmp / f'out{k}'
            t0 = time.perf_counter()
            series.run_timeseries(*paths, d.met[:24 * days], out, d.longitude, d.latitude, utc_offset_hours=0.0,
                                  no_write=not write)
            dt = time.perf_counter() - t0
            shutil.rmtree(out, ignore_errors=True)
            return dt

        res = {}
        for days, write, reps in ((1, True, repeats), (30, False, repeats), (30, True, 1)):
            if reps > 1:
                run(days, write, 0)
            ts = [run(days, write, k + 1) for k in range(reps)]
            key = f'{days}_day{"s" if days > 1 else ""}_{"written" if write else "no_write"}_s'
            res[key] = float(np.median(ts))
            log(f'[bench] e2e run_timeseries 2000x2000, {days} day(s), {"written" if write else "no-write"}: '
                f'{res[key]:.2f} s (runs {", ".join(f"{t:.2f}" for t in ts)})')
        res['reference_published_s'] = {'1_day': 2.7, '30_days': 21.0,
                                        'note': 'RTX 4090, complete workflow incl. GeoTIFF writes (README.md:17-24)'}
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=30, help='timed days (30: the whole 720-hour c2 month)')
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--size', type=int, default=2000)
    ap.add_argument('--eager', action='store_true', help='headline with eager launches instead of the hipGraph')
    ap.add_argument('--eager-days', type=int, default=3, help='days of the eager, event-timed leg (0 = skip)')
    ap.add_argument('--form', choices=('auto', 'pipelined', 'resident', 'hourly'), default='auto',
                    help='day launch form: timed choice, resident day (2 launches), or 2 launches per hour')
    ap.add_argument('--other-days', type=int, default=3,
                    help='days timed in the launch form not chosen for the headline (0 = skip)')
    ap.add_argument('--workers', type=int, default=0, help='table worker processes (0 = host cores - 1, at most 8)')
    ap.add_argument('--no-workers', action='store_true',
                    help='build the day tables in this process (for profiler runs, which must not spawn)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='run a 256 MiB streaming copy (sw_copy) 3 times for HBM counter scaling')
    ap.add_argument('--no-adjoint', action='store_true', help='skip the forward + adjoint leg')
    ap.add_argument('--no-tune', action='store_true', help='keep the default 32x8 tile (no autotuning)')
    ap.add_argument('--tile', type=int, default=0,
                    help='fixed tile width (16, 32, 64, 128) instead of the autotuner (for profiling runs)')
    ap.add_argument('--plain-days', type=int, default=1,
                    help='also time the plain per-step forward for this many days (0 = skip)')
    ap.add_argument('--stepwise-days', type=int, default=1,
                    help='also time the reference-decomposition per-step forward for this many days (0 = skip)')
    ap.add_argument('--no-c4', action='store_true', help='skip the 64-domain uneven batch leg')
    ap.add_argument('--c4-steps', type=int, default=8, help='timed days of the c4 batch (each run starts with a lone shadow pass and ends with a lone pixel pass)')
    ap.add_argument('--c4-count', type=int, default=64, help='domains of the c4 batch')
    ap.add_argument('--c4-audit', type=int, default=3,
                    help='c4 domains audited against the FP64 step decomposition (smallest, median, largest, ...)')
    ap.add_argument('--no-e2e', action='store_true', help='skip the end-to-end run_timeseries leg')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is '
              f'described, not implemented', file=sys.stderr)
        sys.exit(2)
    from solweig_amd import runtime as rt, synthetic, batched
    workers = args.workers or None
    days = 30 if args.size == 2000 else 3
    log(f'[bench] generating c2 morphology {args.size}x{args.size}')
    t0 = time.perf_counter()
    dom = synthetic.make_c1(seed=2, size=args.size, days=days, start=datetime.date(2025, 6, 6), name='c2')
    log(f'[bench] generated in {time.perf_counter() - t0:.1f} s; static preparation')
    static = static_leg(dom)
    t0 = time.perf_counter()
    eng = batched.Engine([dom], n_hours=24)
    rt.synchronize()
    static['engine_setup_s'] = time.perf_counter() - t0
    log(f'[bench] static prep {static["engine_setup_s"]:.2f} s (host derivations, uploads, sky view, sky sums); '
        f'device: walls {static["walls_ms"]:.2f} ms, SVF march {static["svf_visibility_ms"]:.1f} ms, '
        f'SVF reduce {static["svf_reduce_ms"]:.2f} ms, sky sums {static["sky_sums_ms"]:.2f} ms')
    P = eng.domains[0].P
    tuned = None
    if args.tile:
        eng.set_tiles(args.tile)
        tuned = dict(tile_cols=args.tile, fixed=True)
        log(f'[bench] fixed tile width {args.tile}')
    elif not args.no_tune:
        from solweig_amd import autotune
        tuned = autotune.tune(eng, eng.tables(day_index=0), use_cache=False)
        log(f'[bench] autotuned tile width {tuned["tile_cols"]} (ms over hours {tuned.get("hours")}: {tuned["ms"]})')
    eng.workers = n_workers = 0 if args.no_workers else eng.start_table_workers(workers)
    log(f'[bench] {n_workers} table worker processes')
    use_graph = not args.eager
    n_total = args.warmup + args.steps + max(0, args.eager_days) + (2 * (1 + args.other_days) if args.other_days > 0 else 0)
    pipe = DayPipeline(eng, [k % days for k in range(n_total)])
    # Launch form: the first three days, cycled with their uploads, in both forms (resident day and
    # two launches per hour), alternated; the engine keeps the faster (autotune.tune_launch_form).
    # The device also reaches its sustained clocks here after idling while the workers started.
    # Outputs of these days are overwritten by the days that follow.
    form = None
    if args.form == 'auto':
        from solweig_amd import autotune
        t0 = time.perf_counter()
        form = autotune.tune_launch_form(eng, [eng.tables(day_index=k) for k in range(3)], use_graph=use_graph)
        log(f'[bench] launch form: {form["form"]} (ms per day: {form["ms"]}; {time.perf_counter() - t0:.2f} s)')
    else:
        from solweig_amd import autotune
        autotune.set_launch_form(eng, args.form)
    resident = eng.resident
    form_name = 'pipelined' if eng.pipelined else ('resident' if eng.resident else 'hourly')
    run_days(eng, pipe, args.warmup, use_graph)
    rt.synchronize()
    log(f'[bench] {args.warmup} warm-up days done')
    # Timed: the table workers build the next days while the device replays the current one.
    t0 = time.perf_counter()
    run_days(eng, pipe, args.steps, use_graph)
    rt.synchronize()
    dt = time.perf_counter() - t0
    value = P * 24 * args.steps / dt
    log(f'[bench] c2 {"graph" if use_graph else "eager"}: {dt / args.steps * 1e3:.2f} ms/day, {value:.3e} px-h/s')
    tmrt, utci = eng.outputs(0)
    log(f'[bench] Tmrt range {np.nanmin(tmrt):.2f}..{np.nanmax(tmrt):.2f}, UTCI {np.nanmin(utci):.2f}..{np.nanmax(utci):.2f}')
    # the last timed day again, alone (one resident day, no overlap): its outputs must be the bits
    # the timed run left
    last = (args.warmup + args.steps - 1) % days
    keep = eng.resident, eng.pipelined
    eng.resident, eng.pipelined = True, False
    eng.load(eng.tables(day_index=last))
    eng.reset_state()
    eng.forward_day()
    t1, u1 = eng.outputs(0)
    eng.resident, eng.pipelined = keep
    eng.graph = None
    last_day_check = dict(day=int(last), identical_to_the_day_run_alone=bool(
        np.array_equal(tmrt, t1, equal_nan=True) and np.array_equal(utci, u1, equal_nan=True)))
    log(f'[bench] last timed day (day {last}) run alone: identical {last_day_check["identical_to_the_day_run_alone"]}')

    # eager leg: per-pass HIP events on the engine's stream (the kernels' launch durations)
    eager = roof = None
    if args.eager_days > 0:
        evs = [[rt.Event() for _ in range(DIFF_EVENT + 2)] for _ in range(args.eager_days)]
        t0 = time.perf_counter()
        etabs = run_days(eng, pipe, args.eager_days, False, events=evs)
        rt.synchronize()
        edt = time.perf_counter() - t0
        n_day_hours = sum(int(etabs[k][0].hours[h, 0] != 0) for k in range(args.eager_days) for h in range(24))
        if eng.resident:
            # two launches per day: the shadow pass of all hours, the pixel pass of all hours
            t_sh = sum(evs[k][0].elapsed_ms(evs[k][1]) for k in range(args.eager_days))
            t_df = sum(evs[k][1].elapsed_ms(evs[k][DIFF_EVENT]) for k in range(args.eager_days))
            t_px = sum(evs[k][DIFF_EVENT].elapsed_ms(evs[k][2]) for k in range(args.eager_days))
            launches, hours_per_launch = args.eager_days, 24
            names = ('k_hour_shadow (all hours)', 'k_day_pixels')
        else:
            t_sh = t_px = 0.0
            t_df = sum(evs[k][DIFF_EVENT].elapsed_ms(evs[k][DIFF_EVENT + 1]) for k in range(args.eager_days))
            for k in range(args.eager_days):
                for h in range(24):
                    t_sh += evs[k][3 * h].elapsed_ms(evs[k][3 * h + 1])
                    t_px += evs[k][3 * h + 1].elapsed_ms(evs[k][3 * h + 2])
            launches, hours_per_launch = 24 * args.eager_days, 1
            names = ('k_hour_shadow', 'k_hour_pixels')
        sh_ms, px_ms, df_ms = t_sh / launches, t_px / launches, t_df / args.eager_days
        eager = dict(value=P * 24 * args.eager_days / edt, ms_per_day=edt / args.eager_days * 1e3,
                     days=args.eager_days, device_ms_per_day=(t_sh + t_df + t_px) / args.eager_days,
                     avg_ms={names[0]: sh_ms, names[1]: px_ms, 'k_day_diffuse (all hours)': df_ms},
                     hours_per_launch=hours_per_launch, daytime_hours_per_day=n_day_hours / args.eager_days)
        dom_ms, dom_name, dom_b = (px_ms, names[1], B_PIXELS) if px_ms >= sh_ms else (sh_ms, names[0], B_SHADOW)
        dom_bytes = dom_b * P * hours_per_launch
        achieved = dom_bytes / (dom_ms * 1e-3)
        pair = B_ALG * P * hours_per_launch / ((sh_ms + px_ms) * 1e-3)
        roof = dict(bound='hbm', achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s', frac=achieved / HBM_PEAK,
                    traffic=TRAFFIC_2000.get(dom_name) if args.size == 2000 else None,
                    valu_issue=(dict(wave_instructions_per_launch=VALU_2000[dom_name],
                                     achieved=VALU_2000[dom_name] / (dom_ms * 1e-3), peak=VALU_PEAK,
                                     frac=VALU_2000[dom_name] / (dom_ms * 1e-3) / VALU_PEAK, unit='wave-instr/s',
                                     source=f'{TRAFFIC_SOURCE} (SQ_INSTS_VALU, a profile value)')
                                if args.size == 2000 and dom_name in VALU_2000 else None),
                    traffic_source=f'{TRAFFIC_SOURCE} (rocprofv3 FETCH_SIZE x copy calibration + WRITE_SIZE; '
                                   f'a profile value, not measured in this run)',
                    kernel=dom_name, algorithmic_bytes_per_launch=dom_bytes, avg_ms=dom_ms,
                    algorithmic_bytes_per_pixel_hour=dom_b, hours_per_launch=hours_per_launch,
                    pair=dict(kernels=f'{names[0]} + {names[1]}', achieved=pair / 1e9, frac=pair / HBM_PEAK,
                              algorithmic_bytes=B_ALG * P * hours_per_launch, avg_ms=sh_ms + px_ms))
        log(f'[bench] eager: {eager["ms_per_day"]:.2f} ms/day; per launch shadow {sh_ms:.4f} ms, pixels {px_ms:.4f} ms '
            f'({hours_per_launch} h per launch), diffuse sums {df_ms:.4f} ms per day; {dom_name} {achieved / 1e9:.0f} GB/s = '
            f'{100 * achieved / HBM_PEAK:.1f} % of HBM')
    other = None
    if args.other_days > 0:
        # the same days in the launch forms not chosen (autotune.FORMS)
        from solweig_amd import autotune
        chosen = 'pipelined' if eng.pipelined else ('resident' if eng.resident else 'hourly')
        other = {}
        for f in autotune.FORMS:
            if f == chosen:
                continue
            autotune.set_launch_form(eng, f)
            run_days(eng, pipe, 1, use_graph)
            rt.synchronize()
            t0 = time.perf_counter()
            run_days(eng, pipe, args.other_days, use_graph)
            rt.synchronize()
            odt = time.perf_counter() - t0
            ov = P * 24 * args.other_days / odt
            other[f] = dict(value=ov, ms_per_day=odt / args.other_days * 1e3, days=args.other_days,
                            headline_over_this=value / ov)
            log(f'[bench] launch form {f}: {other[f]["ms_per_day"]:.2f} ms/day ({ov:.3e} px-h/s); '
                f'headline / this = {value / ov:.3f}')
        autotune.set_launch_form(eng, chosen)
    eng.close_table_workers()

    plain = None
    if args.plain_days > 0:
        from solweig_amd import model, series
        from solweig_amd.sun import met_sun
        prep = model.prepare_from_domain(eng.domains[0], dom.location, dom.scale, stream=eng.stream)
        mets = [dom.met[k * 24:(k + 1) * 24] for k in range(days)]
        series.plain_day(prep, mets[0], met_sun(mets[0], dom.location, dom.utc),
                         ground=eng.domains[0].dev['ground'], keep=False)
        rt.synchronize()
        t0 = time.perf_counter()
        for k in range(args.plain_days):
            series.plain_day(prep, mets[k + 1], met_sun(mets[k + 1], dom.location, dom.utc),
                             ground=eng.domains[0].dev['ground'], keep=False)
        rt.synchronize()
        pdt = time.perf_counter() - t0
        plain = dict(value=P * 24 * args.plain_days / pdt, ms_per_day=pdt / args.plain_days * 1e3,
                     batched_over_plain=value / (P * 24 * args.plain_days / pdt),
                     note='per-step API: host tables + H2D + full-output shadow and pixel passes + UTCI '
                          'kernel + stream sync every hour, outputs kept on the device')
        log(f'[bench] plain per-step forward {plain["value"]:.3e} px-h/s ({plain["ms_per_day"]:.1f} ms/day)')
        del prep
    if args.stepwise_days > 0:
        from solweig_amd import series
        from solweig_amd.sun import met_sun
        t0 = time.perf_counter()
        ctx = series.stepwise_context(dom, dtype=np.float32)
        rt.synchronize()
        setup = time.perf_counter() - t0
        mets = [dom.met[k * 24:(k + 1) * 24] for k in range(days)]
        series.stepwise_day(*ctx, mets[0], met_sun(mets[0], dom.location, dom.utc))
        rt.synchronize()
        t0 = time.perf_counter()
        for k in range(args.stepwise_days):
            m = mets[(k + 1) % days]
            series.stepwise_day(*ctx, m, met_sun(m, dom.location, dom.utc))
        rt.synchronize()
        sdt = time.perf_counter() - t0
        sv = P * 24 * args.stepwise_days / sdt
        stepwise = dict(value=sv, ms_per_day=sdt / args.stepwise_days * 1e3, batched_over_stepwise=value / sv,
                        setup_s=setup,
                        note='the reference decomposition of the step (model.py:15-256): standalone sub-model '
                             'kernels in FP32 in Solweig_2022a_calc order, a stream sync after each, host '
                             'scalars per hour, UTCI kernel, Tmrt and UTCI copied to the host every hour')
        if plain is None:
            plain = {}
        plain['stepwise'] = stepwise
        log(f'[bench] stepwise (reference decomposition) forward {sv:.3e} px-h/s ({stepwise["ms_per_day"]:.0f} '
            f'ms/day; batched / stepwise = {value / sv:.1f}x; setup {setup:.1f} s)')
        del ctx
    if args.copy_calibration:
        src = rt.DeviceArray((64 << 20,), np.float32)
        dst = rt.DeviceArray((64 << 20,), np.float32)
        for _ in range(3):
            rt.call('sw_copy', dst.ptr, src.ptr, src.nbytes, eng.stream.handle)
        eng.stream.sync()
        log('[bench] copy calibration: 3 x 256 MiB read + 256 MiB written')
    adj = None if args.no_adjoint else adjoint_leg()
    del eng
    c4 = None if args.no_c4 else c4_leg(args, workers)
    e2e = None if args.no_e2e else e2e_leg()
    cpu = None if args.no_cpu else cpu_baseline()
    out = dict(metric='Tmrt+UTCI pixel-hours/s at 2000^2 x24h, 1 GPU; batched forward',
               value=value, unit='pixel-hours/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=dt / args.steps * 1e3, higher_is_better=True, scaling='weak', vs_baseline=None,
               dtype='f32 (f64 masks/lag/UTCI; int8 x int8 -> int32 diffuse sums)', data='synthetic',
               config=dict(workload=f'c2 batched forward {args.size}x{args.size}, 24 h per step (30-day forcing '
                                    f'cycled), day tables built '
                                    f'{f"by {n_workers} CPU worker processes" if n_workers else "in the main process"}'
                                    f' inside the timed region, '
                                    f'{"hipGraph replay" if use_graph else "eager launches"} of '
                                    f'{FORM_TEXT[form_name]}',
                           graph=use_graph, launch_form=form if form is not None else form_name,
                           last_day_check=last_day_check,
                           resident=resident, static_prep=static, tile=tuned,
                           eager=eager, other_launch_form=other,
                           plain_forward=plain,
                           forward_adjoint=adj, c4=c4, end_to_end=e2e),
               roofline=roof, cpu_baseline=cpu)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
