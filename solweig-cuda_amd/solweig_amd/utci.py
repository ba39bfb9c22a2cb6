"""UTCI for the device path (drop-in for utci.utci_calculator, utci.py:6-33).

The day loop evaluates UTCI with scalar weather (series.py:466-467): the polynomial of
Broede et al. (2012) is factored on the host into seven FP64 coefficients of powers of
dTmrt = Tmrt - Ta (the role of utci.cu:243-498), and the device evaluates it per pixel
(sw_utci_scalar, or fused into the hour kernel).
"""
from __future__ import annotations

import numpy as np

from .utci_table import TERMS

_TERMS = np.asarray(TERMS, dtype=np.float64)
_ES = (2.7150305, -2.8365744E3, -6.028076559E3, 1.954263612E1, -2.737830188E-2,
       1.6261698E-5, 7.0229056E-10, -1.8680009E-13)


def vapour_pressure_kpa(Ta, RH):
    """Pa = es(Ta) * RH / 100 / 10 with es in hPa over water (Hardy 1998; utci.cu:249-260)."""
    tk = Ta + 273.15
    c = _ES
    es = c[0] * np.log(tk) + c[1] * tk ** -2.0 + c[2] * tk ** -1.0 + c[3] + c[4] * tk + c[5] * tk ** 2.0 \
        + c[6] * tk ** 3.0 + c[7] * tk ** 4.0
    return np.exp(es) * 0.01 * RH / 100.0 / 10.0


def coefficients(Ta, RH, va):
    """c_0..c_6 (FP64) with UTCI = sum_k c_k (Tmrt - Ta)^k for scalar Ta (C), RH (%), va (m/s)."""
    Ta, RH, va = float(Ta), float(RH), float(va)
    Pa = float(vapour_pressure_kpa(Ta, RH))
    powers = lambda x: x ** np.arange(7, dtype=np.float64)
    ta_p, va_p, pa_p = powers(Ta), powers(va), powers(Pa)
    a = _TERMS[:, 1].astype(int)
    b = _TERMS[:, 2].astype(int)
    d = _TERMS[:, 3].astype(int)
    p = _TERMS[:, 4].astype(int)
    contrib = _TERMS[:, 0] * ta_p[a] * va_p[b] * pa_p[p]
    out = np.zeros(7)
    np.add.at(out, d, contrib)
    out[0] += Ta
    return out


def utci_calculator(Ta, RH, Tmrt, va10m, *, ground=None, stream=None):
    """UTCI of a device Tmrt plane (runtime.DeviceArray, float32) for scalar weather.

    Returns a new DeviceArray. ``ground`` (DeviceArray uint8, optional) NaNs non-ground pixels
    as series.py:468-469 does."""
    from . import runtime as rt
    ta, rh, va = np.float32(Ta), np.float32(RH), np.float32(va10m)
    sentinel = int(ta <= -999 or rh <= -999 or va <= -999)
    c = np.full(7, -999.0) if sentinel else coefficients(float(ta), float(rh), float(va))
    out = rt.DeviceArray(Tmrt.shape, np.float32)
    n = int(np.prod(Tmrt.shape))
    rt.call('sw_utci_scalar', Tmrt.ptr, n, c.ctypes.data, float(ta), sentinel,
            ground.ptr if ground is not None else None, out.ptr, rt._s(stream))
    return out
