"""Benchmark: Tmrt+UTCI pixel-hours/s of the batched day forward on one MI355X.

A step is one day (24 hourly steps) of the batched forward over the 2000x2000 c2 workload
(seeded synthetic morphology, its 30-day forcing cycled): per step the host builds the day's
tables, uploads them and the device runs 24 x (shadow pass + pixel pass); Tmrt and UTCI stay in
device memory. Static preparation (sky view) runs before the timed region and is reported apart.

Prints one JSON line (see the fields below). ``--gpus N`` with N > 1 is refused: this build runs
on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'solweig-cuda_amd'))

import numpy as np  # noqa: E402

B_ALG = 157.5          # algorithmic bytes per pixel-hour of the forward (static inputs + outputs)
HBM_PEAK = 8.0e12      # MI355X HBM3E bytes/s
# HBM bytes of one hour (shadow + pixel launch, mean over a day) of the 2000x2000 workload, from
# rocprofv3 FETCH_SIZE (x 1.9999 copy calibration) + WRITE_SIZE; profiles/k_hour_pair_c2_2000x2000.md
TRAFFIC_PER_HOUR_2000 = 0.977e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(hours=tuple(range(5, 21)), size=200):
    """The FP64 NumPy oracle (a port of the UMEP step) on the 200x200 c0 workload, a bounded
    sample of daytime hours after its static preparation, on one core."""
    os.environ.setdefault('OMP_NUM_THREADS', '1')
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    from solweig_amd import synthetic, sun as psun
    from oracle import day as oday, umep
    d = synthetic.make_c0(size=size)
    s = psun.met_sun(d.met, d.location, d.utc)
    t0 = time.perf_counter()
    p = oday.prepare(d.dsm, d.trees, d.dem, d.walls, d.aspect, d.latitude)
    t_static = time.perf_counter() - t0
    hrs = list(hours)
    met = d.met
    sub = {k: np.asarray(v)[hrs[0]:hrs[-1] + 1] for k, v in s.items()}
    t0 = time.perf_counter()
    oday.run_day(p, met[hrs[0]:hrs[-1] + 1], sub)
    dt = time.perf_counter() - t0
    px_h = size * size * len(hrs)
    return dict(value=px_h / dt, unit='pixel-hours/s', cores=1, kind='port',
                sample=f'oracle FP64 NumPy UMEP step on c0 {size}x{size}, hours {hrs[0]}-{hrs[-1]} '
                       f'({px_h} px-h in {dt:.1f} s; static SVF prep {t_static:.1f} s excluded)')


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
                       stream, events=ev[:3])
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


def c4_main(args):
    """The uneven-size batch: 64 domains (sides U_int[256, 1536], latitudes U(0, 60) N, each with
    its own morphology and forcing), all in one launch per pass and hour. Timed: the device day
    with the day tables resident (one day of forcing, built before the timed region and reported
    separately). Two domains are also run alone and must match their part of the batch bit for bit."""
    from solweig_amd import runtime as rt, synthetic, batched
    t0 = time.perf_counter()
    doms = synthetic.make_c4(count=args.c4_count)
    sides = [d.shape[0] for d in doms]
    P = sum(d.dsm.size for d in doms)
    log(f'[bench] c4: {len(doms)} domains, sides {min(sides)}..{max(sides)}, {P / 1e6:.1f} Mpx '
        f'(generated in {time.perf_counter() - t0:.1f} s)')
    t0 = time.perf_counter()
    eng = batched.Engine(doms, n_hours=24)
    rt.synchronize()
    t_prep = time.perf_counter() - t0
    eng.set_tiles(args.tile or 16)
    log(f'[bench] c4 static preparation {t_prep:.1f} s, {eng.total_blocks} tiles')
    t0 = time.perf_counter()
    tabs = eng.tables()
    t_tab = time.perf_counter() - t0
    eng.load(tabs)
    for _ in range(args.warmup):
        eng.reset_state()
        eng.forward_day()
        eng.stream.sync()
    rt.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.reset_state()
        eng.forward_day()
    eng.stream.sync()
    rt.synchronize()
    dt = time.perf_counter() - t0
    value = P * 24 * args.steps / dt
    log(f'[bench] c4: {dt / args.steps * 1e3:.1f} ms/day, {value:.3e} px-h/s; host tables {t_tab * 1e3:.0f} ms/day')
    checks = []
    for k in (int(np.argmin(sides)), int(np.argmax(sides))):
        tm, ut = eng.outputs(k)
        alone = batched.Engine([doms[k]], n_hours=24)
        alone.set_tiles(args.tile or 16)
        alone.load(alone.tables())
        alone.reset_state()
        alone.forward_day()
        tm1, ut1 = alone.outputs(0)
        same = bool(np.array_equal(tm, tm1, equal_nan=True) and np.array_equal(ut, ut1, equal_nan=True))
        checks.append(dict(domain=k, side=sides[k], identical_to_single_domain_run=same))
        log(f'[bench] c4 domain {k} ({sides[k]}^2) alone vs in the batch: identical {same}')
        del alone
    out = dict(metric='Tmrt+UTCI pixel-hours/s, uneven batch of 64 domains, 1 GPU',
               value=value, unit='pixel-hours/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=dt / args.steps * 1e3, higher_is_better=True, scaling='weak', vs_baseline=None,
               dtype='f32 (f64 masks/lag/wedge/UTCI)', data='synthetic',
               config=dict(workload=f'c4 uneven batch: {len(doms)} domains, sides {min(sides)}..{max(sides)}, '
                                    f'{P} pixels, 24 h per step, one launch per pass and hour, day tables resident',
                           tile=args.tile or 16, static_prep_s=t_prep, host_tables_ms=t_tab * 1e3,
                           checks=checks))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--size', type=int, default=2000)
    ap.add_argument('--graph', action='store_true', help='replay the day as a captured hipGraph')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='run a 256 MiB streaming copy (sw_copy) 3 times for HBM counter scaling')
    ap.add_argument('--no-adjoint', action='store_true', help='skip the forward + adjoint leg')
    ap.add_argument('--no-tune', action='store_true', help='keep the default 32x8 tile (no autotuning)')
    ap.add_argument('--tile', type=int, default=0,
                    help='fixed tile width (16, 32, 64, 128) instead of the autotuner (for profiling runs)')
    ap.add_argument('--plain-days', type=int, default=1,
                    help='also time the plain per-step forward for this many days (0 = skip)')
    ap.add_argument('--workload', choices=('c2', 'c4'), default='c2',
                    help='c2: the 2000x2000 day (default); c4: the 64-domain uneven batch')
    ap.add_argument('--c4-count', type=int, default=64, help='domains of the c4 batch')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is a '
              f'design note, not part of this build', file=sys.stderr)
        sys.exit(2)
    if args.workload == 'c4':
        c4_main(args)
        return
    from solweig_amd import runtime as rt, synthetic, batched
    days = 30 if args.size == 2000 else 3
    log(f'[bench] generating c2 morphology {args.size}x{args.size}')
    t0 = time.perf_counter()
    dom = synthetic.make_c1(seed=2, size=args.size, days=days, start=__import__('datetime').date(2025, 6, 6),
                            name='c2')
    log(f'[bench] generated in {time.perf_counter() - t0:.1f} s; static preparation')
    static = static_leg(dom)
    t0 = time.perf_counter()
    eng = batched.Engine([dom], n_hours=24)
    rt.synchronize()
    t_static = time.perf_counter() - t0
    static['engine_setup_s'] = t_static
    log(f'[bench] static prep {t_static:.2f} s (host derivations, uploads, sky view, sky sums); '
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

    def launch_day(evs):
        """Queue the 24 hours of the loaded day; with events, one per pass boundary."""
        if args.graph:
            eng.forward_day(use_graph=True)
            return
        for h in range(24):
            evs[3 * h].record(eng.stream)
            rt.call('sw_hour_shadow', eng.dom_dev.ptr, eng.block_start.ptr, 1, eng.total_blocks, h, 0,
                    eng.stream.handle)
            evs[3 * h + 1].record(eng.stream)
            rt.call('sw_hour_pixels', eng.dom_dev.ptr, eng.block_start.ptr, 1, eng.total_blocks, h,
                    eng.gvf_off.ptr, eng.gvf_bounds.ptr, eng.first, eng.second, 0, eng.stream.handle)
            evs[3 * h + 2].record(eng.stream)

    warm_ev = [rt.Event() for _ in range(3 * 24 + 1)]
    for k in range(args.warmup):
        eng.load(eng.tables(day_index=k % days))
        eng.reset_state()
        launch_day(warm_ev)
        eng.stream.sync()
        log(f'[bench] warmup day {k} done')
    rt.synchronize()
    # Timed: the host builds day k+1's tables (solar position, Perez, rays, patch tables) while the
    # device runs day k; the uploads are queued in stream order behind day k's launches.
    evs = [[rt.Event() for _ in range(3 * 24 + 1)] for _ in range(args.steps)]
    all_tabs = []
    t0 = time.perf_counter()
    tabs = eng.tables(day_index=args.warmup % days)
    for k in range(args.steps):
        eng.load(tabs)
        eng.reset_state()
        launch_day(evs[k])
        all_tabs.append(tabs)
        if k + 1 < args.steps:
            tabs = eng.tables(day_index=(args.warmup + k + 1) % days)
    eng.stream.sync()
    rt.synchronize()
    dt = time.perf_counter() - t0
    t_shadow = t_pixel = 0.0
    n_day_hours = 0
    if not args.graph:
        for k in range(args.steps):
            per_hour = []
            for h in range(24):
                if all_tabs[k][0].hours[h, 0]:
                    n_day_hours += 1
                a = evs[k][3 * h].elapsed_ms(evs[k][3 * h + 1])
                b = evs[k][3 * h + 1].elapsed_ms(evs[k][3 * h + 2])
                t_shadow += a
                t_pixel += b
                per_hour.append(f'{h}:{a:.2f}/{b:.2f}(alt {all_tabs[k][0].hours[h, 1]:.0f})')
            log(f'[bench] day {k} per hour shadow/pixel ms: ' + ' '.join(per_hour))
    px_h = P * 24 * args.steps
    value = px_h / dt
    launches = 24 * args.steps
    pair_s = (t_shadow + t_pixel) / 1e3 / launches if launches and t_pixel else None
    tmrt, utci = eng.outputs(0)
    log(f'[bench] Tmrt range {np.nanmin(tmrt):.2f}..{np.nanmax(tmrt):.2f}, UTCI {np.nanmin(utci):.2f}..{np.nanmax(utci):.2f}')
    roof = None
    if pair_s:
        achieved = B_ALG * P / pair_s
        roof = dict(bound='hbm', achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s',
                    frac=achieved / HBM_PEAK,
                    traffic=TRAFFIC_PER_HOUR_2000 if args.size == 2000 else None,
                    algorithmic_bytes_per_hour=B_ALG * P,
                    kernel='k_hour_shadow + k_hour_pixels (one hour of the forward)',
                    avg_ms=dict(k_hour_shadow=t_shadow / launches, k_hour_pixels=t_pixel / launches,
                                daytime_hours=n_day_hours / max(args.steps, 1),
                                device_ms_per_day=(t_shadow + t_pixel) / args.steps))
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
    if args.copy_calibration:
        src = rt.DeviceArray((64 << 20,), np.float32)
        dst = rt.DeviceArray((64 << 20,), np.float32)
        for _ in range(3):
            rt.call('sw_copy', dst.ptr, src.ptr, src.nbytes, eng.stream.handle)
        eng.stream.sync()
        log('[bench] copy calibration: 3 x 256 MiB read + 256 MiB written')
    adj = None if args.no_adjoint else adjoint_leg()
    cpu = None if args.no_cpu else cpu_baseline()
    out = dict(metric='Tmrt+UTCI pixel-hours/s at 2000^2 x24h, 1 GPU; batched forward',
               value=value, unit='pixel-hours/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=dt / args.steps * 1e3, higher_is_better=True, scaling='weak', vs_baseline=None,
               dtype='f32 (f64 masks/lag/wedge/UTCI)', data='synthetic',
               config=dict(workload=f'c2 batched forward {args.size}x{args.size}, 24 h per step '
                                    f'(30-day forcing cycled), host day tables inside the timed region '
                                    f'(built for day k+1 while the device runs day k)',
                           graph=bool(args.graph), static_prep=static, tile=tuned, plain_forward=plain,
                           forward_adjoint=adj),
               roofline=roof, cpu_baseline=cpu)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
