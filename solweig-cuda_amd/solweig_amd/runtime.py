"""ctypes binding of libsolweig_hip.so (the C ABI declared in include/solweig_hip.h).

This is the only place that talks to the device: device buffers, streams, events, graphs and the
kernel entry points. There is no CPU fallback - if the library cannot be loaded every call raises.
Replaces the reference's CuPy/NVRTC runtime (runtime.py:11-76).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, 'libsolweig_hip.so')

SW_HOUR_STRIDE = 64
SW_PATCH_F = 12
SW_PATCH_D = 4
SW_PATCHES = 153
SW_VIS_WORDS = 5
SW_MAX_PIXELS = (1 << 30) - 1   # largest raster of one domain (32-bit byte offsets of the marches)
SWS_NIB, SWS_CSUN, SWS_CSH, SWS_MASK, SWS_ELIG, SW_SEARCH_WORDS = 0, 1248, 1760, 2272, 2536, 2560
# the day's diffuse table of sw_day_diffuse: int8 digits in fragment order, then
# double scale [rows], two rows per hour, rows = diff_rows(n_hours)
SW_DIFF_K, SW_DIFF_DIGITS = 320, 4


def diff_rows(n_hours):
    return 16 * ((int(n_hours) + 7) // 8)


def diff_table_bytes(n_hours):
    return diff_rows(n_hours) * (SW_DIFF_DIGITS * SW_DIFF_K + 8)
TILE_X, TILE_Y = 32, 8          # default workgroup tile (256 threads)
TILE_CHOICES = (16, 32, 64, 128)  # tile widths the hour kernels accept (height = 256 / width)


def tile_count(shape, tile_cols=TILE_X):
    rows, cols = shape
    ty = 256 // tile_cols
    return ((cols + tile_cols - 1) // tile_cols) * ((rows + ty - 1) // ty)

# hour record fields (enum order of solweig_hip.h)
_HOUR_FIELDS = ('DAY ALTITUDE AZIMUTH TANZEN TA RADG RADD RADI PSI SINALT COSALT ALTMAX FRAC SNUP '
                'CITGG LAG_WEIGHT LAG_FIRST LAG_UPDATE FACE_LOW FACE_HIGH FACE_MIDDLE RAY_START RAY_LEN '
                'LWALL FLUX4 LW_AMBIENT LW_SUNLIT LW_REFL SW_SUNLIT SW_SHADED TA32 UTCI0').split()
H = {name: i for i, name in enumerate(_HOUR_FIELDS)}
H['PATCH_TABLE'] = H['UTCI0'] + 7
for _k, _name in enumerate('TGOUT_WEIGHT TGOUT_UPDATE ALBEDO_B ABSK ABSL FUP FSIDE FCYL DIFF_VEG TGWALL '
                           'SKY_A SKY_B ELIG'.split()):
    H[_name] = H['PATCH_TABLE'] + 1 + _k
H['TG_SINE'], H['TG_TMAX'] = H['ELIG'] + 5, H['ELIG'] + 6
PF = {n: i for i, n in enumerate('LV KD LDN LSD WCOS WSIN DE DW DN DS ELIGIBLE'.split())}
PD = {n: i for i, n in enumerate('YI SUN_BELOW SHADE_FROM'.split())}
SKY_SUMS = 10
SW_EXTRA_PLANES = 5
EXTRA_NAMES = ('Kdown', 'Kup', 'Ldown', 'Lup', 'Shadow')   # enum SWE_* of solweig_hip.h
FULL_PLANES = ('Kdown Kup Ldown Lup Tg Keast Ksouth Kwest Knorth Least Lsouth Lwest Lnorth KsideI '
               'TgOut Lside KsideD dRad Kside shadow wallsun F_sh').split()


class SwDomain(ctypes.Structure):
    """Mirror of sw_domain: every field 8 bytes."""
    _fields_ = [(n, ctypes.c_int64) for n in ('rows', 'cols', 'pixels', 'landcover')] + \
        [('dsm_raise', ctypes.c_double), ('tile_cols', ctypes.c_int64)] + \
        [(n, ctypes.c_uint64) for n in ('dsm', 'trees', 'walls', 'aspect', 'emis', 'alb', 'tgk', 'tstart',
                                        'tmaxlst', 'svfbuveg', 'tan_asvf', 'tan_svfalfa', 'ground', 'vis',
                                        'rec', 'lag', 'tmrt', 'utci', 'hours', 'ptab_f', 'ptab_d',
                                        'ray_off', 'ray_dz', 'masks', 'shadow_planes', 'full', 'skysums',
                                        'hmax')] + [('hmax_cols', ctypes.c_int64)] + \
        [(n, ctypes.c_uint64) for n in ('gvf_len', 'gvf_nosh', 'march_trees', 'ptab_s', 'gvf_edge',
                                        'diffuse_tab', 'diffuse')]
HMAX_CELL = 16


class SwAdjoint(ctypes.Structure):
    """Mirror of sw_adjoint: device pointers of the adjoint buffers."""
    _fields_ = [(n, ctypes.c_uint64) for n in ('lam_state', 'lam_rec', 'coef', 'grad_alb', 'grad_emis',
                                               'lam_patch', 'svfveg', 'scalars', 'patch', 'objective',
                                               'partial')] + \
        [('reserved', ctypes.c_int64 * 5)]


# slots of the per-hour adjoint record (enum SWA_* of solweig_hip.h)
_ADJ_FIELDS = 'TA PSI CITGG LWALL ALBEDO_B RADI RADD RADG FLUX4 SW_SUNLIT SW_SHADED SKY_A SKY_B LW_AMBIENT LW_SUNLIT TA32 UTCI0'.split()
ADJ = {name: i for i, name in enumerate(_ADJ_FIELDS)}
ADJ['TRANS_STATIC'] = ADJ['UTCI0'] + 7
ADJ_SLOTS = 24
ADJ_PARTIAL_SLOTS = ADJ_SLOTS + 2
ADJ_PATCH = {'SKY_KDOWN': 0, 'VEG_KDOWN': 1, 'SKY_KSIDE': 2, 'VEG_KSIDE': 3}


class SolweigHipError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded library; raises if it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise SolweigHipError(f'{LIBRARY} is not built; run __graft_entry__.build()')
    L = ctypes.CDLL(LIBRARY)
    vp, i32, i64, f64, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_size_t
    sig = {
        'sw_version': ([], i32), 'sw_last_error': ([], ctypes.c_char_p), 'sw_record_bytes': ([i32], i32),
        'sw_device_count': ([ctypes.POINTER(i32)], i32), 'sw_set_device': ([i32], i32),
        'sw_malloc': ([ctypes.POINTER(vp), sz], i32), 'sw_free': ([vp], i32),
        'sw_host_alloc': ([ctypes.POINTER(vp), sz], i32), 'sw_host_free': ([vp], i32),
        'sw_event_sync': ([vp], i32),
        'sw_memcpy_h2d': ([vp, vp, sz, vp], i32), 'sw_memcpy_d2h': ([vp, vp, sz, vp], i32),
        'sw_memcpy_d2d': ([vp, vp, sz, vp], i32), 'sw_memset': ([vp, i32, sz, vp], i32),
        'sw_memcpy_d2h_async': ([vp, vp, sz, vp], i32), 'sw_stream_wait_event': ([vp, vp], i32),
        'sw_stream_create': ([ctypes.POINTER(vp)], i32), 'sw_stream_destroy': ([vp], i32),
        'sw_stream_sync': ([vp], i32), 'sw_event_create': ([ctypes.POINTER(vp)], i32),
        'sw_event_destroy': ([vp], i32), 'sw_event_record': ([vp, vp], i32),
        'sw_event_elapsed_ms': ([vp, vp, ctypes.POINTER(ctypes.c_float)], i32),
        'sw_device_sync': ([], i32),
        'sw_svf_visibility': ([vp, vp, vp, vp, i32, f64, i32, i32, vp, vp, vp, i32, i32, vp, vp], i32),
        'sw_svf_reduce': ([vp, vp, vp, f64, i32, i32, vp, vp, vp, i32, i32, f64, vp, vp], i32),
        'sw_hour_shadow': ([vp, vp, i32, i32, i32, i32, vp], i32),
        'sw_hour_pixels': ([vp, vp, i32, i32, i32, vp, vp, i32, i32, i32, vp], i32),
        'sw_hours_forward': ([vp, vp, i32, i32, i32, i32, vp, vp, i32, i32, vp], i32),
        'sw_day_forward': ([vp, vp, i32, i32, i32, i32, vp, vp, i32, i32, vp], i32),
        'sw_day_shadow': ([vp, vp, i32, i32, i32, i32, vp], i32),
        'sw_day_pixels': ([vp, vp, i32, i32, i32, i32, vp, vp, i32, i32, vp], i32),
        'sw_day_pixels_capped': ([vp, vp, i32, i32, i32, i32, vp, vp, i32, i32, i32, vp], i32),
        'sw_day_diffuse': ([vp, vp, i32, i32, i32, vp], i32),
        'sw_utci_scalar': ([vp, i64, vp, f64, i32, vp, vp, vp], i32),
        'sw_sun_position': ([vp, i32, i32, i32, f64, f64, f64, f64, vp, vp], i32),
        'sw_copy': ([vp, vp, sz, vp], i32),
        'sw_sky_sums': ([vp, i64, vp, vp, vp], i32),
        'sw_adjoint_hour': ([vp, vp, i32, i32, i32, vp, vp, vp, i32, i32, f64, vp], i32),
        'sw_day_adjoint': ([vp, vp, i32, i32, i32, vp, vp, vp, i32, i32, f64, vp], i32),
        'sw_graph_begin': ([vp], i32), 'sw_graph_end': ([vp, ctypes.POINTER(vp)], i32),
        'sw_graph_launch': ([vp, vp], i32), 'sw_graph_destroy': ([vp], i32),
        'sw_gvf_static': ([vp, vp, vp, vp, i32, i32, vp, vp, i32, i32, vp, vp, vp, vp], i32),
        'sw_gvf_fixed_steps': ([vp, i32], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    for kind, rec in ((0, SwDomain), (1, SwAdjoint)):
        if L.sw_record_bytes(kind) != ctypes.sizeof(rec):
            raise SolweigHipError(f'{LIBRARY} was built for {L.sw_record_bytes(kind)}-byte {rec.__name__} records, '
                                  f'these bindings use {ctypes.sizeof(rec)}; rebuild it')
    _lib = L
    return L


EXPORTED = ('sw_version', 'sw_record_bytes', 'sw_last_error', 'sw_device_count', 'sw_set_device', 'sw_malloc', 'sw_free',
            'sw_host_alloc', 'sw_host_free', 'sw_event_sync',
            'sw_memcpy_h2d', 'sw_memcpy_d2h', 'sw_memcpy_d2d', 'sw_memset', 'sw_stream_create',
            'sw_stream_destroy', 'sw_stream_sync', 'sw_event_create', 'sw_event_destroy',
            'sw_event_record', 'sw_event_elapsed_ms', 'sw_device_sync', 'sw_svf_visibility',
            'sw_svf_reduce', 'sw_sky_sums', 'sw_hour_shadow', 'sw_hour_pixels', 'sw_hours_forward', 'sw_day_forward',
            'sw_day_shadow', 'sw_day_pixels', 'sw_day_pixels_capped', 'sw_day_diffuse', 'sw_sun_position',
            'sw_utci_scalar',
            'sw_copy', 'sw_graph_begin', 'sw_graph_end', 'sw_graph_launch', 'sw_graph_destroy',
            'sw_adjoint_hour', 'sw_day_adjoint', 'sw_cylindric_wedge', 'sw_wave_delay', 'sw_kup', 'sw_lside',
            'sw_surface_temperature', 'sw_assemble_tmrt', 'sw_gvf', 'sw_pack_visibility', 'sw_longwave',
            'sw_kside', 'sw_dsm_walls', 'sw_memcpy_d2h_async', 'sw_stream_wait_event', 'sw_gvf_static',
            'sw_gvf_fixed_steps', 'sw_gvf_noalloc', 'sw_step_terms', 'sw_diffuse')


def check(rc):
    if rc != 0:
        msg = lib().sw_last_error()
        raise SolweigHipError(f'libsolweig_hip error {rc}: {msg.decode() if msg else ""}')


def call(name, *args):
    check(getattr(lib(), name)(*args))


class DeviceView:
    """A typed window [offset, offset + shape) into another device allocation."""

    def __init__(self, owner, offset_bytes, shape, dtype):
        self.owner, self.offset = owner, int(offset_bytes)
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize

    @property
    def ptr(self):
        return self.owner.ptr + self.offset

    def to_host(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        call('sw_memcpy_d2h', out.ctypes.data, self.ptr, self.nbytes, _s(stream))
        return out

    def get(self):
        return self.to_host()


class DeviceArray:
    """A device allocation with a NumPy dtype and shape (row-major)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if np.ndim(shape) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ptr = ctypes.c_void_p()
        call('sw_malloc', ctypes.byref(self._ptr), self.nbytes)

    @property
    def ptr(self):
        return self._ptr.value or 0

    @classmethod
    def from_host(cls, array, dtype=None, stream=None):
        a = np.ascontiguousarray(array, dtype=dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, array, stream=None):
        a = np.ascontiguousarray(array, dtype=self.dtype)
        if a.nbytes > self.nbytes:
            raise ValueError('upload larger than the allocation')
        call('sw_memcpy_h2d', self.ptr, a.ctypes.data, a.nbytes, _s(stream))
        call('sw_stream_sync', _s(stream))  # the host array may be pageable and short-lived

    def to_host(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        call('sw_memcpy_d2h', out.ctypes.data, self.ptr, self.nbytes, _s(stream))
        return out

    def zero(self, stream=None):
        call('sw_memset', self.ptr, 0, self.nbytes, _s(stream))

    def get(self):
        return self.to_host()

    def plane(self, k, shape=None):
        """View of plane k of a [planes, ...] array, optionally reshaped."""
        inner = self.shape[1:]
        n = int(np.prod(inner, dtype=np.int64)) * self.dtype.itemsize
        return DeviceView(self, k * n, shape or inner, self.dtype)

    def free(self):
        if self._ptr.value:
            lib().sw_free(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory (sw_host_alloc) exposed as a NumPy array, for uploads that run
    asynchronously in stream order. The caller must not overwrite it while a copy from it is
    still pending (wait on an event recorded after that copy)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if np.ndim(shape) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ptr = ctypes.c_void_p()
        call('sw_host_alloc', ctypes.byref(self._ptr), max(self.nbytes, 1))
        buf = (ctypes.c_char * max(self.nbytes, 1)).from_address(self._ptr.value)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(self.shape))).reshape(self.shape)

    @property
    def ptr(self):
        return self._ptr.value or 0

    def __del__(self):
        try:
            if self._ptr.value:
                self.array = None
                lib().sw_host_free(self._ptr)
                self._ptr = ctypes.c_void_p()
        except Exception:
            pass


def _s(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class Stream:
    def __init__(self):
        self._h = ctypes.c_void_p()
        call('sw_stream_create', ctypes.byref(self._h))

    @property
    def handle(self):
        return self._h.value

    def sync(self):
        call('sw_stream_sync', self.handle)

    def wait(self, event):
        """Later work on this stream waits for ``event`` (recorded on any stream)."""
        call('sw_stream_wait_event', self.handle, event.handle)

    def __del__(self):
        try:
            if self._h.value:
                lib().sw_stream_destroy(self._h)
        except Exception:
            pass


class Event:
    def __init__(self):
        self._h = ctypes.c_void_p()
        call('sw_event_create', ctypes.byref(self._h))

    @property
    def handle(self):
        return self._h.value

    def record(self, stream=None):
        call('sw_event_record', self._h, _s(stream))

    def sync(self):
        call('sw_event_sync', self._h)

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        call('sw_event_elapsed_ms', self._h, later._h, ctypes.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            if self._h.value:
                lib().sw_event_destroy(self._h)
        except Exception:
            pass


class Graph:
    """A captured launch sequence (hipGraphExec)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def capture(cls, stream, body):
        call('sw_graph_begin', _s(stream))
        try:
            body()
        finally:
            h = ctypes.c_void_p()
            call('sw_graph_end', _s(stream), ctypes.byref(h))
        return cls(h)

    def launch(self, stream):
        call('sw_graph_launch', self._h, _s(stream))

    def __del__(self):
        try:
            if self._h.value:
                lib().sw_graph_destroy(self._h)
        except Exception:
            pass


def device_count():
    n = ctypes.c_int()
    check(lib().sw_device_count(ctypes.byref(n)))
    return n.value


def synchronize():
    call('sw_device_sync')
