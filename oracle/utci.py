"""UTCI in FP64 (oracle; test infrastructure only).

The polynomial of Broede et al. (2012) as expanded in the reference's fused kernel
(utci.cu:1-241), its per-hour factorisation in powers of dTmrt (utci.cu:243-498), and the
scalar-weather evaluation used by the day loop (utci.py:22-33; series.py:466-469).
"""
from __future__ import annotations

import importlib.util
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location(
    '_oracle_utci_table', os.path.join(_HERE, '..', 'solweig-cuda_amd', 'solweig_amd', 'utci_table.py'))
_tab = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_tab)
TERMS = np.asarray(_tab.TERMS, dtype=np.float64)

_ES = (2.7150305, -2.8365744E3, -6.028076559E3, 1.954263612E1, -2.737830188E-2,
       1.6261698E-5, 7.0229056E-10, -1.8680009E-13)


def saturation_hpa(Ta):
    """Saturation vapour pressure over water in hPa (Hardy 1998 ITS-90; utci.cu:13-22)."""
    tk = np.asarray(Ta, dtype=np.float64) + 273.15
    c = _ES
    es = c[0] * np.log(tk) + c[1] * tk ** -2.0 + c[2] * tk ** -1.0 + c[3] + c[4] * tk \
        + c[5] * tk ** 2.0 + c[6] * tk ** 3.0 + c[7] * tk ** 4.0
    return np.exp(es) * 0.01


def utci(Ta, RH, Tmrt, va):
    """UTCI in C for air temperature Ta (C), RH (%), Tmrt (C), 10 m wind va (m/s); -999 sentinel."""
    Ta, RH, Tmrt, va = np.broadcast_arrays(*(np.asarray(x, dtype=np.float64) for x in (Ta, RH, Tmrt, va)))
    Pa = saturation_hpa(Ta) * RH / 100.0 / 10.0
    D = Tmrt - Ta
    out = Ta.copy()
    for c, a, b, d, p in TERMS:
        out = out + c * Ta ** a * va ** b * D ** d * Pa ** p
    bad = (Ta <= -999) | (RH <= -999) | (va <= -999) | (Tmrt <= -999)
    return np.where(bad, -999.0, out)


def coefficients(Ta, RH, va):
    """c_0..c_6 with UTCI = sum_k c_k dTmrt^k for scalar weather (c_0 includes Ta)."""
    Ta, RH, va = float(Ta), float(RH), float(va)
    Pa = float(saturation_hpa(Ta)) * RH / 100.0 / 10.0
    c = np.zeros(7)
    c[0] = Ta
    for coef, a, b, d, p in TERMS:
        c[int(d)] += coef * Ta ** a * va ** b * Pa ** p
    return c
