"""Ground view factors of one step as a standalone device call (model.py:512-622).

``prepare_gvf_geometry`` and ``gvf`` keep the reference's names, arguments and results:
``gvf`` returns a GvfResult whose ``outputs`` are the 17 planes of gvf_2018a (gvfLup, gvfalb,
gvfalbnosh, the E, S, W, N triples, gvfSum, gvfNorm) and whose ``surface_temperature`` is Tg with
water pixels (lc_grid == 3) replaced by Twater - Ta, as sunonsurface_2018a leaves it
(umep_solweig.py:1419-1421). It runs the sw_gvf kernels of libsolweig_hip in the requested
precision for any caller fields (fractional shadow, a ground mask of any value). The fused hour
pass computes the same sums inside its pixel kernel and does not call this.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import runtime as rt
from . import sky
from .flux import field_table, working_dtype

_SIG = set()


def _entry(name='sw_gvf'):
    f = getattr(rt.lib(), name)
    if name not in _SIG:
        vp, ci = ctypes.c_void_p, ctypes.c_int
        work = [vp] if name == 'sw_gvf' else []
        f.argtypes = [vp, vp, ci, ci, ci, vp, vp, ci, ci, ci, vp, ci, ci] + work + [vp, vp]
        f.restype = ci
        _SIG.add(name)
    return f


@dataclass(frozen=True)
class GvfGeometry:
    """Static state of ``gvf`` (model.py:512-524): FP64 walls, ground mask and wall aspect in
    radians on the device, the 18 x second_steps walk offsets and the facing bounds."""
    walls: object
    buildings: object
    aspect: object
    offsets: object
    shade_bounds: object
    shape: tuple
    first_steps: int
    second_steps: int
    scale: float
    step_stride: int


@dataclass(frozen=True)
class GvfResult:
    outputs: tuple
    surface_temperature: object


def _host(v):
    return v.get() if hasattr(v, 'get') else v


def prepare_gvf_geometry(walls, buildings, dirwalls, scale, first, second):
    """model.py:533-566: walls, buildings (ground mask) and dirwalls (degrees) of the raster."""
    scale, first, second = float(scale), float(first), float(second)
    walls = np.ascontiguousarray(np.asarray(_host(walls), dtype=np.float64))
    shape = walls.shape
    buildings = np.ascontiguousarray(np.broadcast_to(np.asarray(_host(buildings), dtype=np.float64), shape))
    aspect = np.broadcast_to(np.asarray(_host(dirwalls), dtype=np.float64), shape) * np.pi / 180
    off, bounds, f_steps, s_steps = sky.gvf_geometry(scale, first, second)
    if s_steps < 1:
        raise ValueError('prepare_gvf_geometry: round(second * scale) must be at least 1')
    up = rt.DeviceArray.from_host
    return GvfGeometry(up(walls), up(buildings), up(np.ascontiguousarray(aspect)), up(off), up(bounds),
                       tuple(int(x) for x in shape), int(f_steps), int(s_steps), scale, int(off.shape[1]))


def gvf(geometry, wallsun, shadow, Tg, Tgwall, Ta, emis_grid, ewall, alb_grid, SBC, albedo_b, Twater=None,
        lc_grid=None, landcover=0, *, dtype=np.float64, no_alloc=False):
    """Ground view factors of one step (model.py:569-622) -> GvfResult. ``no_alloc``: the walk's
    samples go to the library's static scratch (sw_gvf_noalloc, rasters up to 2097152 pixels in
    FP32, 1048576 in FP64) instead of a work buffer allocated for the call."""
    g = geometry
    dtype = working_dtype(dtype)
    landcover = int(float(landcover))
    if landcover == 1:
        if Twater is None or lc_grid is None:
            raise ValueError('gvf: landcover 1 needs Twater and lc_grid')
        twater = float(Twater)
    else:
        twater, lc_grid = 0.0, 0.0
    rows, cols = g.shape
    n = rows * cols
    values = (g.buildings, g.walls, g.aspect, wallsun, shadow, Tg, Tgwall, emis_grid, alb_grid, lc_grid)
    P, C, keep = field_table(values, n, dtype, fp64_fields=(0, 1, 2, 3, 9))
    out = rt.DeviceArray((18, rows, cols), dtype)
    params = (float(Ta), float(ewall), float(SBC), float(albedo_b), float(landcover), twater)
    Q = (ctypes.c_double * 6)(*params)
    head = (P, C, len(values), rows, cols, g.offsets.ptr, g.shade_bounds.ptr, g.step_stride, g.first_steps,
            g.second_steps, Q, 6, dtype.itemsize)
    if no_alloc:
        work = None
        rt.check(_entry('sw_gvf_noalloc')(*head, out.ptr, None))
    else:
        work = rt.DeviceArray((n, 8), dtype)
        rt.check(_entry()(*head, work.ptr, out.ptr, None))
    rt.check(rt.lib().sw_stream_sync(None))
    del keep, work
    return GvfResult(tuple(out.plane(k, g.shape) for k in range(17)), out.plane(17, g.shape))
