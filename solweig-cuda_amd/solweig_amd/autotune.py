"""Tile-width autotuner of the hour kernels.

Every hour kernel maps a 256-thread workgroup to a tile of W x (256 / W) pixels, W in
runtime.TILE_CHOICES; a wave then covers 64 / W rows of W pixels. The width trades the
coalescing of the neighbour gathers (wide rows) against the L1/L2 reuse of the walks and rays
between the rows of a wave (tall tiles). ``tune`` times the shadow + pixel pass of representative
hours of the loaded day (the lowest and the highest daytime sun, where the ray march and the
ground-view walk respectively dominate) for every width on the actual domains, keeps the fastest,
and remembers it per domain shape in a small JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import runtime as rt

CACHE = os.environ.get('SOLWEIG_AMD_TUNE_CACHE',
                       os.path.join(os.path.expanduser('~'), '.cache', 'solweig_amd', 'tiles.json'))


def _key(engine):
    return ';'.join(f'{d.shape[0]}x{d.shape[1]}' for d in engine.domains) + f'|h{engine.n_hours}'


def _load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key, value):
    data = _load_cache()
    data[key] = value
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, 'w') as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass


def representative_hours(tables):
    """Daytime hours with the lowest and the highest sun of the first domain's day."""
    h = tables[0].hours
    day = np.flatnonzero(h[:, rt.H['DAY']] != 0)
    if len(day) == 0:
        return []
    alt = h[day, rt.H['ALTITUDE']]
    return sorted({int(day[np.argmin(alt)]), int(day[np.argmax(alt)])})


def tune(engine, tables=None, candidates=rt.TILE_CHOICES, repeats=3, use_cache=True):
    """Pick the tile width of ``engine`` (all domains alike); returns {'tile_cols', 'ms'}.
    The day's tables are loaded into the engine; the lag state is cleared afterwards."""
    key = _key(engine)
    if use_cache:
        hit = _load_cache().get(key)
        if hit and hit.get('tile_cols') in candidates:
            engine.set_tiles(hit['tile_cols'])
            return dict(hit, cached=True)
    tables = tables if tables is not None else engine.tables(day_index=0)
    engine.load(tables)
    hours = representative_hours(tables)
    if not hours:
        return dict(tile_cols=engine.domains[0].tile_cols, ms={}, cached=False)
    ev = [rt.Event() for _ in range(2)]
    times = {}
    for w in candidates:
        engine.set_tiles(w)
        engine.reset_state()
        for h in hours:                                  # warm-up of this launch geometry
            engine.hour(h)
        best = float('inf')
        for _ in range(repeats):
            ev[0].record(engine.stream)
            for h in hours:
                engine.hour(h)
            ev[1].record(engine.stream)
            engine.stream.sync()
            best = min(best, ev[0].elapsed_ms(ev[1]))
        times[int(w)] = best
    choice = min(times, key=times.get)
    engine.set_tiles(choice)
    engine.reset_state()
    result = dict(tile_cols=int(choice), ms=times, hours=hours)
    if use_cache:
        _store_cache(key, result)
    return dict(result, cached=False)


FORMS = ('pipelined', 'resident', 'hourly')


def set_launch_form(engine, form):
    """Set ``engine``'s launch form: 'pipelined' (resident days whose shadow pass overlaps the
    previous day's pixel pass: Engine.run_days), 'resident' (sw_day_forward) or 'hourly'
    (sw_hours_forward)."""
    if form not in FORMS:
        raise ValueError(f'launch form {form!r} not in {FORMS}')
    engine.resident, engine.pipelined, engine.graph = form != 'hourly', form == 'pipelined', None


def tune_launch_form(engine, tables=None, days=10, rounds=3, use_graph=True, forms=FORMS):
    """Pick how the engine launches its days: 'pipelined' (resident days, each day's all-hours
    shadow pass running on a side stream alongside the previous day's pixel pass, Engine.run_days),
    'resident' (sw_day_forward: the shadow pass of every hour in one launch, then one pixel pass
    with the lag state in registers, one day after the other) or 'hourly' (two launches per hour,
    sw_hours_forward). ``tables``: a list of day tables (one day's list per entry) cycled with an
    upload before every timed day, as a run of consecutive days does (default: the loaded day,
    replayed). Times ``days`` consecutive days in each form, alternating the forms for ``rounds``
    rounds (so that clock and cache drift falls on all), keeps the fastest by the best round and
    leaves it set on the engine; returns {'form', 'resident', 'pipelined', 'ms'} (ms per day)."""
    cycle = list(tables) if tables is not None else [None]
    if cycle == [None] and 'pipelined' in forms:
        raise ValueError('the pipelined form needs the day tables')
    ev = [rt.Event() for _ in range(2)]
    best = {f: float('inf') for f in forms}

    def day(k):
        if cycle[k % len(cycle)] is not None:
            engine.load(cycle[k % len(cycle)])
        engine.reset_state()
        engine.forward_day(use_graph=use_graph)

    for _ in range(rounds):
        for form in forms:
            set_launch_form(engine, form)
            if form == 'pipelined':
                engine.run_days([cycle[0], cycle[1 % len(cycle)]], use_graph=use_graph)   # capture / warm
                ev[0].record(engine.stream)
                engine.run_days((cycle[(k + 1) % len(cycle)] for k in range(days)), use_graph=use_graph)
            else:
                day(0)                                      # capture / warm this form
                ev[0].record(engine.stream)
                for k in range(days):
                    day(k + 1)
            ev[1].record(engine.stream)
            engine.stream.sync()
            best[form] = min(best[form], ev[0].elapsed_ms(ev[1]) / days)
    form = min(forms, key=lambda f: best[f])
    set_launch_form(engine, form)
    engine.reset_state()
    return dict(form=form, resident=engine.resident, pipelined=engine.pipelined, ms=best)
