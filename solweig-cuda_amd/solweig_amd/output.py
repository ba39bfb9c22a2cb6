"""Output path of the day loop (the reference's output.py:11-216 and runtime.py:79-163).

Each finished day's Tmrt and UTCI are copied, in stream order on the engine's stream, into one
of ``slots`` device staging buffers; a side stream waits for that copy and moves the buffer into
page-locked host memory; a writer thread waits for the host copy and hands the arrays to the
caller's ``write`` callback, then frees the slot. The engine's stream never waits for the host:
the next day's launches queue behind the staging copy only, so the device-to-host transfer and
the file writes overlap the following days. ``submit`` blocks only when every slot is still
being written (the bounded queue of the reference's output worker).
"""
from __future__ import annotations

import queue
import threading

import numpy as np

from . import runtime as rt


class DayOutputs:
    """``write(tag, tmrt, utci)``, or ``write(tag, tmrt, utci, extras)`` with extras a dict of the
    engine's extra planes ({name: [n_hours, rows, cols]}) when the engine keeps them."""

    def __init__(self, engine, k=0, slots=2):
        self.engine, self.k = engine, int(k)
        d = engine.domains[self.k]
        self.extra = 'extra' in engine.buf[self.k]
        self.shape = (2, engine.n_hours, *d.shape)
        self.extra_shape = (engine.n_hours, rt.SW_EXTRA_PLANES, *d.shape)
        n = engine.n_hours * d.P
        self._bytes = n * 4
        self._extra_bytes = rt.SW_EXTRA_PLANES * n * 4 if self.extra else 0
        words = 2 * n + (rt.SW_EXTRA_PLANES * n if self.extra else 0)
        self._dev = [rt.DeviceArray((words,), np.float32) for _ in range(slots)]
        self._host = [rt.PinnedArray((words,), np.float32) for _ in range(slots)]
        self._staged = [rt.Event() for _ in range(slots)]
        self._landed = [rt.Event() for _ in range(slots)]
        self._copy = rt.Stream()
        self._free = queue.Queue()
        for s in range(slots):
            self._free.put(s)
        self._jobs = queue.Queue()
        self._error = None
        self._thread = threading.Thread(target=self._run, name='solweig-output', daemon=True)
        self._thread.start()

    def submit(self, tag, write):
        """Queue the outputs of the day just enqueued on the engine's stream; ``write(tag, tmrt,
        utci)`` runs on the writer thread with host arrays [n_hours, rows, cols] that are valid
        only during the call."""
        self._raise()
        slot = self._free.get()
        b = self.engine.buf[self.k]
        st, dev = self.engine.stream, self._dev[slot]
        rt.call('sw_memcpy_d2d', dev.ptr, b['tmrt'].ptr, self._bytes, st.handle)
        rt.call('sw_memcpy_d2d', dev.ptr + self._bytes, b['utci'].ptr, self._bytes, st.handle)
        if self.extra:
            rt.call('sw_memcpy_d2d', dev.ptr + 2 * self._bytes, b['extra'].ptr, self._extra_bytes, st.handle)
        self._staged[slot].record(st)
        self._copy.wait(self._staged[slot])
        rt.call('sw_memcpy_d2h_async', self._host[slot].ptr, dev.ptr, 2 * self._bytes + self._extra_bytes,
                self._copy.handle)
        self._landed[slot].record(self._copy)
        self._jobs.put((slot, tag, write))

    def _run(self):
        while True:
            job = self._jobs.get()
            if job is None:
                return
            slot, tag, write = job
            try:
                self._landed[slot].sync()
                flat = self._host[slot].array
                a = flat[:2 * (self._bytes // 4)].reshape(self.shape)
                if self._error is None:
                    if self.extra:
                        e = flat[2 * (self._bytes // 4):].reshape(self.extra_shape)
                        write(tag, a[0], a[1], {name: e[:, i] for i, name in enumerate(rt.EXTRA_NAMES)})
                    else:
                        write(tag, a[0], a[1])
            except BaseException as e:          # reported by the next submit or close
                self._error = e
            finally:
                self._free.put(slot)

    def _raise(self):
        if self._error is not None:
            e, self._error = self._error, None
            raise e

    def close(self):
        """Wait until every queued day is written; re-raises a writer error."""
        self._jobs.put(None)
        self._thread.join()
        self._raise()
