"""FP64 NumPy restatement of the SOLWEIG 2022a step (oracle; test infrastructure only).

Follows umep_solweig.py (UMEP 4.1.1 as retained by SOLWEIG-CUDA) for the branches that
series.py exercises: cylinder body (cyl=1), anisotropic sky (patch option 2, 153 patches),
vegetation scheme, global radiation split by Reindl. Array operations keep UMEP's order so that
masks compare exactly and fluxes agree to FP64 rounding.
"""
from __future__ import annotations

import math

import numpy as np

SBC = 5.67051e-8
DEG = np.pi / 180


# --------------------------------------------------------------------------- host scalars

def daylen(doy, lat):
    """Sunrise/sunset hours (umep_solweig.py:430-448)."""
    rad = np.pi / 180.0
    dec = -23.45 * np.cos(2.0 * np.pi * (doy + 10.0) / 365.0)
    soc = np.tan(rad * dec) * np.tan(rad * lat)
    soc = min(max(soc, -1.0), 1.0)
    dayl = 12.0 + 24.0 * np.arcsin(soc) / np.pi
    return dayl, dec, 12.0 + dayl / 2.0, 12.0 - dayl / 2.0


def sun_distance(jday):
    """Relative Earth-Sun distance (umep_solweig.py:2513-2523)."""
    b = 2. * np.pi * jday / 365.
    return np.sqrt((1.00011 + np.dot(0.034221, np.cos(b)) + np.dot(0.001280, np.sin(b))
                    + np.dot(0.000719, np.cos((2. * b))) + np.dot(0.000077, np.sin((2. * b)))))


_G_BANDS = ((10., [3.37, 2.85, 2.80, 2.64]), (20., [2.99, 3.02, 2.70, 2.93]),
            (30., [3.60, 3.00, 2.98, 2.93]), (40., [3.04, 3.11, 2.92, 2.94]),
            (50., [2.70, 2.95, 2.77, 2.71]), (60., [2.52, 3.07, 2.67, 2.93]),
            (70., [1.76, 2.69, 2.61, 2.61]), (80., [1.60, 1.67, 2.24, 2.63]),
            (90., [1.11, 1.44, 1.94, 2.02]))


def clearness_index(zen, jday, Ta, RH, radG, lat, P):
    """Crawford-Duchon clearness with the low-sun correction (umep_solweig.py:1654-1734).

    RH is a fraction. Returns (I0, CI, Kt, I0et, CIuncorr)."""
    p = 1013. if P == -999.0 else P * 10.
    Itoa = 1370.0
    D = sun_distance(jday)
    m = 35. * np.cos(zen) * ((1224. * (np.cos(zen) ** 2) + 1) ** (-1 / 2.))
    Trpg = 1.021 - 0.084 * (m * (0.000949 * p + 0.051)) ** 0.5
    G = None
    lower = -np.inf
    for upper, g in _G_BANDS:
        if (lat < upper) and (lat >= lower if lower > -np.inf else True):
            G = g
            break
        lower = upper
    if jday > 335 or jday <= 60:
        G = G[0]
    elif 60 < jday <= 152:
        G = G[1]
    elif 152 < jday <= 244:
        G = G[2]
    else:
        G = G[3]
    a2, b2 = 17.27, 237.7
    Td = (b2 * (((a2 * Ta) / (b2 + Ta)) + np.log(RH))) / (a2 - (((a2 * Ta) / (b2 + Ta)) + np.log(RH)))
    Td = (Td * 1.8) + 32
    u = np.exp(0.1133 - np.log(G + 1) + 0.0393 * Td)
    Tw = 1 - 0.077 * ((u * m) ** 0.3)
    Tar = 0.935 ** m
    I0 = Itoa * np.cos(zen) * Trpg * Tw * D * Tar
    if abs(zen) > np.pi / 2:
        I0 = 0
    if not np.isreal(I0):
        I0 = 0
    corr = 0.1473 * np.log(90 - (zen / np.pi * 180)) + 0.3454
    with np.errstate(divide='ignore', invalid='ignore'):
        CIuncorr = np.float64(radG) / I0
    CI = CIuncorr + (1 - corr)
    I0et = Itoa * np.cos(zen) * D
    Kt = radG / I0et
    if math.isnan(CI):
        CI = float('Inf')
    return I0, CI, Kt, I0et, CIuncorr


def diffuse_fraction(radG, altitude, Kt, Ta, RH):
    """Reindl et al. (1990) direct/diffuse split (umep_solweig.py:1789-1832). RH in %."""
    alfa = altitude * (np.pi / 180)
    if Ta <= -999.00 or RH <= -999.00 or np.isnan(Ta) or np.isnan(RH):
        if Kt <= 0.3:
            radD = radG * (1.020 - 0.248 * Kt)
        elif 0.3 < Kt < 0.78:
            radD = radG * (1.45 - 1.67 * Kt)
        else:
            radD = radG * 0.147
    else:
        RH = RH / 100
        if Kt <= 0.3:
            radD = radG * (1 - 0.232 * Kt + 0.0239 * np.sin(alfa) - 0.000682 * Ta + 0.0195 * RH)
        elif 0.3 < Kt < 0.78:
            radD = radG * (1.329 - 1.716 * Kt + 0.267 * np.sin(alfa) - 0.00357 * Ta + 0.106 * RH)
        else:
            radD = radG * (0.426 * Kt - 0.256 * np.sin(alfa) + 0.00349 * Ta + 0.0734 * RH)
    radI = (radG - radD) / (np.sin(alfa))
    if radI < 0:
        radI = 0
    if altitude < 1 and radI > radG:
        radI = radG
    if radD > radG:
        radD = radG
    return radI, radD


_BANDS = {1: (30, 30, 24, 24, 18, 12, 6, 1), 2: (31, 30, 28, 24, 19, 13, 7, 1),
          3: (62, 60, 56, 48, 38, 26, 14, 1)}


def create_patches(option=2):
    """Sky-vault patch centres in degrees (umep_solweig.py:1737-1786)."""
    alt_int = np.array([6, 18, 30, 42, 54, 66, 78, 90])
    start = np.array([0, 4, 2, 5, 8, 0, 10, 0])
    counts = np.array(_BANDS[option])
    alt, azi = [], []
    for j in range(len(alt_int)):
        step = 360 / counts[j]
        for k in range(counts[j]):
            alt.append(float(alt_int[j]))
            azi.append(k * step + start[j])
    return np.array(alt, dtype=np.float64), np.array(azi, dtype=np.float64)


_PEREZ = np.array([
    [1.3525, -1.2219, -1.1000, -0.5484, -0.6000, -1.0156, -1.0000, -1.0500],
    [-0.2576, -0.7730, -0.2515, -0.6654, -0.3566, -0.3670, 0.0211, 0.0289],
    [-0.2690, 1.4148, 0.8952, -0.2672, -2.5000, 1.0078, 0.5025, 0.4260],
    [-1.4366, 1.1016, 0.0156, 0.7117, 2.3250, 1.4051, -0.5119, 0.3590],
    [-0.7670, -0.2054, 0.2782, 0.7234, 0.2937, 0.2875, -0.3000, -0.3250],
    [0.0007, 0.0367, -0.1812, -0.6219, 0.0496, -0.5328, 0.1922, 0.1156],
    [1.2734, -3.9128, -4.5000, -5.6812, -5.6812, -3.8500, 0.7023, 0.7781],
    [-0.1233, 0.9156, 1.1766, 2.6297, 1.8415, 3.3750, -1.6317, 0.0025],
    [2.8000, 6.9750, 24.7219, 33.3389, 21.0000, 14.0000, 19.0000, 31.0625],
    [0.6004, 0.1774, -13.0812, -18.3000, -4.7656, -0.9999, -5.0000, -14.5000],
    [1.2375, 6.4477, -37.7000, -62.2500, -21.5906, -7.1406, 1.2438, -46.1148],
    [1.0000, -0.1239, 34.8438, 52.0781, 7.2492, 7.5469, -1.9094, 55.3750],
    [1.8734, -1.5798, -5.0000, -3.5000, -3.5000, -3.4000, -4.0000, -7.2312],
    [0.6297, -0.5081, 1.5218, 0.0016, -0.1554, -0.1078, 0.0250, 0.4050],
    [0.9738, -1.7812, 3.9229, 1.1477, 1.4062, -1.0750, 0.3844, 13.3500],
    [0.2809, 0.1080, -2.6204, 0.1062, 0.3988, 1.5702, 0.2656, 0.6234],
    [0.0356, 0.2624, -0.0156, 0.4659, 0.0032, -0.0672, 1.0468, 1.5000],
    [-0.1246, 0.0672, 0.1597, -0.3296, 0.0766, 0.4016, -0.3788, -0.6426],
    [-0.5718, -0.2190, 0.4199, -0.0876, -0.0656, 0.3017, -2.4517, 1.8564],
    [0.9938, -0.4285, -0.5562, -0.0329, -0.1294, -0.4844, 1.4656, 0.5636]])
"""Perez et al. (1993) luminance coefficients; rows a1..a4, b1..b4, c1..c4, d1..d4, e1..e4,
columns the 8 clearness bins (model.py:810-829 / umep_solweig.py:1924-1944)."""


def perez(zen_deg, azimuth, radD, radI, jday, patch_option=2, clip=False):
    """Relative sky luminance of the 153 patches (umep_solweig.py:1836-2057).

    ``clip`` applies the [-1, 1] clip of the CUDA path's copy (model.py:933). Returns lv [n, 3]
    with columns (altitude deg, azimuth deg, normalised luminance)."""
    coef = _PEREZ
    deg2rad = np.pi / 180
    rad2deg = 180 / np.pi
    altitude = 90 - zen_deg
    zen = zen_deg * deg2rad
    azimuth = azimuth * deg2rad
    altitude = altitude * deg2rad
    Idh, Ibn = radD, radI
    clear = ((Idh + Ibn) / (Idh + 1.041 * np.power(zen, 3))) / (1 + 1.041 * np.power(zen, 3))
    day_angle = jday * 2 * np.pi / 365
    I0 = 1367 * (1.00011 + 0.034221 * np.cos(day_angle) + 0.00128 * np.sin(day_angle)
                 + 0.000719 * np.cos(2 * day_angle) + 0.000077 * np.sin(2 * day_angle))
    if altitude >= 10 * deg2rad:
        airmass = 1 / np.sin(altitude)
    elif altitude < 0:
        airmass = 1 / np.sin(altitude) + 0.50572 * np.power(180 * complex(altitude) / np.pi + 6.07995, -1.6364)
    else:
        airmass = 1 / np.sin(altitude) + 0.50572 * np.power(180 * altitude / np.pi + 6.07995, -1.6364)
    bright = (airmass * radD) / I0
    if Idh <= 10:
        bright = 0
    edges = (1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200)
    k = None
    if clear < edges[0]:
        k = 0
    for j in range(6):
        if edges[j] < clear < edges[j + 1]:
            k = j + 1
    if clear > edges[6]:
        k = 7

    def c(row, col):
        return coef[row * 4 + col, k]
    m_a = c(0, 0) + c(0, 1) * zen + bright * (c(0, 2) + c(0, 3) * zen)
    m_b = c(1, 0) + c(1, 1) * zen + bright * (c(1, 2) + c(1, 3) * zen)
    m_e = c(4, 0) + c(4, 1) * zen + bright * (c(4, 2) + c(4, 3) * zen)
    if k > 0:
        m_c = c(2, 0) + c(2, 1) * zen + bright * (c(2, 2) + c(2, 3) * zen)
        m_d = c(3, 0) + c(3, 1) * zen + bright * (c(3, 2) + c(3, 3) * zen)
    else:
        m_c = np.exp(np.power(bright * (c(2, 0) + c(2, 1) * zen), c(2, 2))) - 1
        m_d = -np.exp(bright * (c(3, 0) + c(3, 1) * zen)) + c(3, 2) + bright * c(3, 3) * bright
    alt, azi = create_patches(patch_option)
    vzen = (90 - alt) * deg2rad
    valt = alt * deg2rad
    vazi = azi * deg2rad
    cos_gamma = np.sin(valt) * np.sin(altitude) + np.cos(altitude) * np.cos(valt) * np.cos(np.abs(vazi - azimuth))
    if clip:
        cos_gamma = np.clip(cos_gamma, -1.0, 1.0)
    lv = (1 + m_a * np.exp(m_b / np.cos(vzen))) * ((1 + m_c * np.exp(m_d * np.arccos(cos_gamma))
                                                      + m_e * cos_gamma * cos_gamma))
    lv = lv / np.sum(lv)
    return np.column_stack((valt * rad2deg, vazi * rad2deg, lv)), clear, bright


def wave_delay(current, firstdaytime, timeadd, timestepdec, state):
    """Surface temperature lag (umep_solweig.py:1633-1652). Returns (Lup, timeadd, state)."""
    if firstdaytime == 1:
        state = current
    if timeadd >= (59 / 1440):
        w = np.exp(-33.27 * timeadd)
        state = current * (1 - w) + state * w
        out = state
        timeadd = timestepdec if timestepdec > (59 / 1440) else 0
    else:
        timeadd = timeadd + timestepdec
        w = np.exp(-33.27 * timeadd)
        out = current * (1 - w) + state * w
    return out, timeadd, state


def cylindric_wedge(zen, svfalfa):
    """Sunlit fraction of walls seen by a cylinder (umep_solweig.py:383-426)."""
    beta, alfa = zen, svfalfa
    with np.errstate(divide='ignore', invalid='ignore'):
        xa = 1 - 2. / (np.tan(alfa) * np.tan(beta))
        ha = 2. / (np.tan(alfa) * np.tan(beta))
        ba = (1. / np.tan(alfa))
        hkil = 2. * ba * ha
        neg = xa < 0
        qa = np.zeros_like(alfa)
        qa[neg] = np.tan(beta) / 2
        Za = np.zeros_like(alfa)
        Za[neg] = ((ba[neg] ** 2) - ((qa[neg] ** 2) / 4)) ** 0.5
        phi = np.zeros_like(alfa)
        phi[neg] = np.arctan(Za[neg] / qa[neg])
        A = np.zeros_like(alfa)
        A[neg] = (np.sin(phi[neg]) - phi[neg] * np.cos(phi[neg])) / (1 - np.cos(phi[neg]))
        ukil = np.zeros_like(alfa)
        ukil[neg] = 2 * ba[neg] * xa[neg] * A[neg]
        Ssurf = hkil + ukil
        return (2 * np.pi * ba - Ssurf) / (2 * np.pi * ba)


# --------------------------------------------------------------------------- shadows

def _shift_windows(dx, dy, sizex, sizey):
    absdx, absdy = abs(dx), abs(dy)
    xc1 = int((dx + absdx) / 2)
    xc2 = int(sizex + (dx - absdx) / 2)
    yc1 = int((dy + absdy) / 2)
    yc2 = int(sizey + (dy - absdy) / 2)
    xp1 = -int((dx - absdx) / 2)
    xp2 = int(sizex - (dx + absdx) / 2)
    yp1 = -int((dy - absdy) / 2)
    yp2 = int(sizey - (dy + absdy) / 2)
    return (slice(xp1, xp2), slice(yp1, yp2)), (slice(xc1, xc2), slice(yc1, yc2))


def march_offsets(azimuth_rad, index, tanazimuth, signsin, signcos):
    """(dx, dy) for ray step ``index`` (umep_solweig.py:2286-2294)."""
    pi4 = np.pi / 4
    if (pi4 <= azimuth_rad < 3 * pi4) or (5 * pi4 <= azimuth_rad < 7 * pi4):
        dy = signsin * index
        dx = -1 * signcos * np.abs(np.round(index / tanazimuth))
        ydom = True
    else:
        dy = signsin * np.abs(np.round(index * tanazimuth))
        dx = -1 * signcos * index
        ydom = False
    return dx, dy, ydom


def shadows(a, vegdem, vegdem2, azimuth, altitude, scale, amaxvalue, bush, walls, aspect):
    """Building and vegetation shadows with wall sun/shade heights
    (shadowingfunction_wallheight_23, umep_solweig.py:2201-2406).

    Returns (vegsh, sh, vbshvegsh, wallsh, wallsun, wallshve, facesh, facesun)."""
    azimuth = azimuth * DEG
    altitude = altitude * DEG
    sizex, sizey = a.shape
    dx = dy = dz = 0
    temp = np.zeros((sizex, sizey))
    tv = np.zeros((sizex, sizey))
    tv2 = np.zeros((sizex, sizey))
    tlf = np.zeros((sizex, sizey))
    tlg = np.zeros((sizex, sizey))
    bushplant = bush > 1
    sh = np.zeros((sizex, sizey))
    vbshvegsh = np.copy(sh)
    vegsh = np.add(np.zeros((sizex, sizey)), bushplant, dtype=float)
    f = np.copy(a)
    shvoveg = np.copy(vegdem)
    wallbol = (walls > 0).astype(float)
    sinaz, cosaz, tanaz = np.sin(azimuth), np.cos(azimuth), np.tan(azimuth)
    ssin, scos = np.sign(sinaz), np.sign(cosaz)
    dssin, dscos = np.abs(1 / sinaz), np.abs(1 / cosaz)
    tanalt = np.tan(altitude) / scale
    index = 0
    dzprev = 0
    while (amaxvalue >= dz) and (np.abs(dx) < sizex) and (np.abs(dy) < sizey):
        dx, dy, ydom = march_offsets(azimuth, index, tanaz, ssin, scos)
        ds = dssin if ydom else dscos
        dz = (ds * index) * tanalt
        tv[:] = 0
        tv2[:] = 0
        temp[:] = 0
        tlf[:] = 0.
        tlg[:] = 0.
        dst, src = _shift_windows(dx, dy, sizex, sizey)
        tv[dst] = vegdem[src] - dz
        tv2[dst] = vegdem2[src] - dz
        temp[dst] = a[src] - dz
        f = np.fmax(f, temp)
        shvoveg = np.fmax(shvoveg, tv)
        sh[f > a] = 1
        sh[f <= a] = 0
        fabovea = (tv > a).astype(int)
        gabovea = (tv2 > a).astype(int)
        tlf[dst] = vegdem[src] - dzprev
        tlg[dst] = vegdem2[src] - dzprev
        lastf = tlf > a
        lastg = tlg > a
        dzprev = dz
        vegsh2 = np.add(np.add(np.add(fabovea, gabovea, dtype=float), lastf, dtype=float), lastg, dtype=float)
        vegsh2[vegsh2 == 4] = 0.
        vegsh2[vegsh2 > 0] = 1.
        vegsh = np.fmax(vegsh, vegsh2)
        vegsh[vegsh * sh > 0] = 0
        vbshvegsh = np.copy(vegsh) + vbshvegsh
        index += 1
    facesh = face_shade(azimuth, aspect, wallbol)
    sh = 1 - sh
    vbshvegsh[vbshvegsh > 0] = 1
    vbshvegsh = vbshvegsh - vegsh
    vegsh[vegsh > 0] = 1
    shvoveg = (shvoveg - a) * vegsh
    vegsh = 1 - vegsh
    vbshvegsh = 1 - vbshvegsh
    shvo = f - a
    facesun = np.logical_and(facesh + (walls > 0).astype(float) == 1, walls > 0).astype(float)
    wallsun = np.copy(walls - shvo)
    wallsun[wallsun < 0] = 0
    wallsun[facesh == 1] = 0
    wallsh = np.copy(walls - wallsun)
    wallshve = shvoveg * wallbol
    wallshve = wallshve - wallsh
    wallshve[wallshve < 0] = 0
    idx = np.where(wallshve > walls)
    wallshve[idx] = walls[idx]
    wallsun = wallsun - wallshve
    idx = np.where(wallsun < 0)
    wallshve[idx] = 0
    wallsun[idx] = 0
    return vegsh, sh, vbshvegsh, wallsh, wallsun, wallshve, facesh, facesun


def face_shade(azimuth_rad, aspect, wallbol):
    """Walls facing away from the sun (umep_solweig.py:2362-2372, 1554-1563)."""
    azilow = azimuth_rad - np.pi / 2
    azihigh = azimuth_rad + np.pi / 2
    if azilow >= 0 and azihigh < 2 * np.pi:
        return np.logical_or(aspect < azilow, aspect >= azihigh).astype(float) - wallbol + 1
    if azilow < 0 and azihigh <= 2 * np.pi:
        azilow = azilow + 2 * np.pi
        return np.logical_or(aspect > azilow, aspect <= azihigh) * -1 + 1
    azihigh = azihigh - 2 * np.pi
    return np.logical_or(aspect > azilow, aspect <= azihigh) * -1 + 1


# --------------------------------------------------------------------------- ground view

def sun_on_surface(azimuthA, scale, buildings, shadow, sunwall, first, second, aspect, walls, Tg,
                   Tgwall, Ta, emis_grid, ewall, alb_grid, albedo_b, Twater, lc_grid, landcover):
    """One GVF search direction (sunonsurface_2018a, umep_solweig.py:1400-1600). Mutates Tg
    where lc_grid == 3 like the original. Returns (gvf, gvfLup, gvfalb, gvfalbnosh, gvf2)."""
    sizex, sizey = walls.shape
    wallbol = (walls > 0) * 1
    sunwall[sunwall > 0] = 1
    azimuth = azimuthA * (np.pi / 180)
    index = 0
    f = buildings
    Lup = SBC * emis_grid * (Tg * shadow + Ta + 273.15) ** 4 - SBC * emis_grid * (Ta + 273.15) ** 4
    if landcover == 1:
        Tg[lc_grid == 3] = Twater - Ta
    Lwall = SBC * ewall * (Tgwall + Ta + 273.15) ** 4 - SBC * ewall * (Ta + 273.15) ** 4
    albshadow = alb_grid * shadow
    alb = alb_grid
    z = lambda: np.zeros((sizex, sizey))
    tempsh, tempbu, tempbub, tempbubwall, tempwallsun = z(), z(), z(), z(), z()
    wsh, wwall, wLupsh, wLwall, walbsh, walbwall, walbnosh, walbwallnosh = z(), z(), z(), z(), z(), z(), z(), z()
    tempLupsh, tempalbsh, tempalbnosh = z(), z(), z()
    first = np.round(first * scale)
    if first < 1:
        first = 1
    second = np.round(second * scale)
    tan_a, ssin, scos = np.tan(azimuth), np.sign(np.sin(azimuth)), np.sign(np.cos(azimuth))
    snap = {}
    for n in np.arange(0, second):
        pi4 = np.pi / 4
        if (pi4 <= azimuth < 3 * pi4) or (5 * pi4 <= azimuth < 7 * pi4):
            dy = ssin * index
            dx = -1 * scos * np.abs(np.round(index / tan_a))
        else:
            dy = ssin * abs(round(index * tan_a))
            dx = -1 * scos * index
        dst, src = _shift_windows(dx, dy, sizex, sizey)
        tempbu[dst] = buildings[src]
        tempsh[dst] = shadow[src]
        tempLupsh[dst] = Lup[src]
        tempalbsh[dst] = albshadow[src]
        tempalbnosh[dst] = alb[src]
        f = np.min([f, tempbu], axis=0)
        wsh = wsh + tempsh * f
        wLupsh = wLupsh + tempLupsh * f
        walbsh = walbsh + tempalbsh * f
        walbnosh = walbnosh + tempalbnosh * f
        tempwallsun[dst] = sunwall[src]
        tempb = tempwallsun * f
        tempbwall = f * -1 + 1
        tempbub = ((tempb + tempbub) > 0) * 1
        tempbubwall = ((tempbwall + tempbubwall) > 0) * 1
        wLwall = wLwall + tempbub * Lwall
        walbwall = walbwall + tempbub * albedo_b
        wwall = wwall + tempbub
        walbwallnosh = walbwallnosh + tempbubwall * albedo_b
        if (n + 1) <= first:
            snap = dict(wall=wwall / 1, sh=wsh / 1, Lwall=wLwall / 1, Lupsh=wLupsh / 1,
                        albwall=walbwall / 1, albsh=walbsh / 1, albwallnosh=walbwallnosh / 1,
                        albnosh=walbnosh / 1)
            snap['sunin'] = snap['wall'] > 0
            snap['wallin'] = snap['albwallnosh'] > 0
        index += 1
    sunin2 = wwall > 0
    wallin2 = walbwallnosh > 0
    facesh = face_shade(azimuth, aspect, wallbol)
    keep = (wwall == second) - facesh
    keep[keep == -1] = 0
    s1 = snap
    gvf1 = ((s1['wall'] + s1['sh']) / (first + 1)) * s1['sunin'] + (s1['sh']) / first * (s1['sunin'] * -1 + 1)
    wwall[keep == 1] = 0
    gvf2 = ((wwall + wsh) / (second + 1)) * sunin2 + wsh / second * (sunin2 * -1 + 1)
    gvf2[gvf2 > 1.] = 1.
    gvfLup1 = ((s1['Lwall'] + s1['Lupsh']) / (first + 1)) * s1['sunin'] + s1['Lupsh'] / first * (s1['sunin'] * -1 + 1)
    wLwall[keep == 1] = 0
    gvfLup2 = ((wLwall + wLupsh) / (second + 1)) * sunin2 + wLupsh / second * (sunin2 * -1 + 1)
    gvfalb1 = ((s1['albwall'] + s1['albsh']) / (first + 1)) * s1['sunin'] + s1['albsh'] / first * (s1['sunin'] * -1 + 1)
    walbwall[keep == 1] = 0
    gvfalb2 = ((walbwall + walbsh) / (second + 1)) * sunin2 + walbsh / second * (sunin2 * -1 + 1)
    gvfalbnosh1 = ((s1['albwallnosh'] + s1['albnosh']) / (first + 1)) * s1['wallin'] \
        + s1['albnosh'] / first * (s1['wallin'] * -1 + 1)
    gvfalbnosh2 = ((walbwallnosh + walbnosh) / second) * wallin2 + walbnosh / second * (wallin2 * -1 + 1)
    gvf = (gvf1 * 0.5 + gvf2 * 0.4) / 0.9
    gvfLup = (gvfLup1 * 0.5 + gvfLup2 * 0.4) / 0.9
    gvfLup = gvfLup + ((SBC * emis_grid * (Tg * shadow + Ta + 273.15) ** 4)
                       - SBC * emis_grid * (Ta + 273.15) ** 4) * (buildings * -1 + 1)
    gvfalb = (gvfalb1 * 0.5 + gvfalb2 * 0.4) / 0.9
    gvfalb = gvfalb + alb_grid * (buildings * -1 + 1) * shadow
    gvfalbnosh = (gvfalbnosh1 * 0.5 + gvfalbnosh2 * 0.4) / 0.9
    gvfalbnosh = gvfalbnosh * buildings + alb_grid * (buildings * -1 + 1)
    return gvf, gvfLup, gvfalb, gvfalbnosh, gvf2


def gvf(wallsun, walls, buildings, scale, shadow, first, second, dirwalls, Tg, Tgwall, Ta, emis_grid,
        ewall, alb_grid, albedo_b, Twater, lc_grid, landcover):
    """18-direction ground view factors (gvf_2018a, umep_solweig.py:583-663).

    Returns the 17-tuple (gvfLup, gvfalb, gvfalbnosh, then E, S, W, N triples, gvfSum, gvfNorm)."""
    directions = np.arange(5, 359, 20)
    shape = walls.shape
    acc = {k: np.zeros(shape) for k in ('Lup', 'alb', 'nosh', 'sum')}
    card = {d: {k: np.zeros(shape) for k in ('Lup', 'alb', 'nosh')} for d in 'ESWN'}
    with np.errstate(divide='ignore', invalid='ignore'):
        sunwall = (wallsun / walls * buildings) == 1
    for az in directions:
        _, lup_i, alb_i, nosh_i, g2 = sun_on_surface(az, scale, buildings, shadow, sunwall, first, second,
                                                      dirwalls * np.pi / 180, walls, Tg, Tgwall, Ta,
                                                      emis_grid, ewall, alb_grid, albedo_b, Twater,
                                                      lc_grid, landcover)
        acc['Lup'] = acc['Lup'] + lup_i
        acc['alb'] = acc['alb'] + alb_i
        acc['nosh'] = acc['nosh'] + nosh_i
        acc['sum'] = acc['sum'] + g2
        for d, ok in (('E', 0 <= az < 180), ('S', 90 <= az < 270), ('W', 180 <= az < 360),
                      ('N', az >= 270 or az < 90)):
            if ok:
                card[d]['Lup'] = card[d]['Lup'] + lup_i
                card[d]['alb'] = card[d]['alb'] + alb_i
                card[d]['nosh'] = card[d]['nosh'] + nosh_i
    n = len(directions)
    ambient = SBC * emis_grid * (Ta + 273.15) ** 4
    out = [acc['Lup'] / n + ambient, acc['alb'] / n, acc['nosh'] / n]
    for d in 'ESWN':
        out += [card[d]['Lup'] / (n / 2) + ambient, card[d]['alb'] / (n / 2), card[d]['nosh'] / (n / 2)]
    norm = acc['sum'] / n
    norm[buildings == 0] = 1
    return tuple(out) + (acc['sum'], norm)


# --------------------------------------------------------------------------- anisotropic sky

def steradians(patch_alt):
    """Solid angle per patch (umep_solweig.py:749-759, 982-990)."""
    skyalt, skyalt_c = np.unique(patch_alt, return_counts=True)
    ster = np.zeros(patch_alt.shape[0])
    for i in range(patch_alt.shape[0]):
        c = skyalt_c[skyalt == patch_alt[i]]
        if c > 1:
            ster[i] = (((360 / c) * DEG) * (np.sin((patch_alt[i] + patch_alt[0]) * DEG)
                                            - np.sin((patch_alt[i] - patch_alt[0]) * DEG)))[0]
        else:
            ster[i] = (((360 / c) * DEG) * (np.sin(patch_alt[i] * DEG)
                                            - np.sin((patch_alt[i - 1] + patch_alt[0]) * DEG)))[0]
    return ster


def shaded_or_sunlit(solar_altitude, solar_azimuth, patch_altitude, patch_azimuth, asvf):
    """Wall patch sunlit/shaded from the SVF horizon angle (umep_solweig.py:1362-1397)."""
    xi = np.cos(np.abs(solar_azimuth - patch_azimuth) * (np.pi / 180))
    yi = 2 * xi * np.tan(solar_altitude * (np.pi / 180))
    hsvf = np.tan(asvf)
    yi_ = 0 if yi > 0 else yi
    deg = np.arctan(hsvf + yi_) * (180 / np.pi)
    return deg < patch_altitude, deg > patch_altitude


def kside_cylinder(radI, radD, shadow, azimuth, altitude, albedo, KupE, KupS, KupW, KupN, lv,
                   diffsh, asvf, shmat, vegshmat, vbshvegshmat):
    """Kside_veg_v2022a with cyl=1 and anisotropic diffuse (umep_solweig.py:666-837).

    Returns (Keast, Ksouth, Kwest, Knorth, KsideI, KsideD, Kside)."""
    shape = shadow.shape
    KsideI = shadow * radI * np.cos(altitude * DEG)
    alt, azi, lum = lv[:, 0], lv[:, 1], lv[:, 2]
    ster = steradians(alt)
    radTot = np.zeros(1)
    for i in range(alt.shape[0]):
        radTot += lum[i] * ster[i] * np.sin(alt[i] * DEG)
    lumChi = (lum * radD) / radTot
    KsideD = np.zeros(shape)
    Kref_sun = np.zeros(shape)
    Kref_sh = np.zeros(shape)
    Kref_veg = np.zeros(shape)
    sunlit_surface = ((albedo * (radI * np.cos(altitude * DEG)) + (radD * 0.5)) / np.pi)
    shaded_surface = ((albedo * radD * 0.5) / np.pi)
    for k in range(alt.shape[0]):
        ang = np.cos(alt[k] * DEG) * np.cos(0)
        KsideD += diffsh[:, :, k] * lumChi[k] * ang * ster[k]
        vegsh = (vegshmat[:, :, k] == 0) | (vbshvegshmat[:, :, k] == 0)
        Kref_veg += shaded_surface * vegsh * ster[k] * np.cos(alt[k] * DEG)
        wall = ((1 - shmat[:, :, k]) * vbshvegshmat[:, :, k]) == 1
        sun, shd = shaded_or_sunlit(altitude, azimuth, alt[k], azi[k], asvf)
        Kref_sun += sunlit_surface * sun * wall * ster[k] * np.cos(alt[k] * DEG)
        Kref_sh += shaded_surface * shd * wall * ster[k] * np.cos(alt[k] * DEG)
    Kside = KsideI + KsideD + Kref_sun + Kref_sh + Kref_veg
    return KupE * 0.5, KupS * 0.5, KupW * 0.5, KupN * 0.5, KsideI, KsideD, Kside


def _kvikt(svf, svfveg, total=4.4897):
    """Building and vegetation view weights of a vertical face (Kvikt_veg, umep_solweig.py:930-943)."""
    poly = lambda x: 63.227 * x ** 6 - 161.51 * x ** 5 + 156.91 * x ** 4 - 70.424 * x ** 3 + 16.773 * x ** 2 - 0.4863 * x
    wall = (total - poly(svf)) / total
    veg = (total - poly(svfveg + svf - 1)) / total
    return veg - wall, wall


def kside(radI, radD, radG, shadow, svfS, svfW, svfN, svfE, svfEveg, svfSveg, svfWveg, svfNveg, azimuth,
          altitude, psi, t, albedo, F_sh, KupE, KupS, KupW, KupN, cyl, lv, anisotropic_diffuse, diffsh, asvf,
          shmat, vegshmat, vbshvegshmat):
    """Kside_veg_v2022a, every branch (umep_solweig.py:666-911): cylinder or box body, anisotropic
    (Perez patches) or weighted isotropic diffuse. Returns (Keast, Ksouth, Kwest, Knorth, KsideI,
    KsideD, Kside); the box and isotropic branches leave KsideD and Kside zero as UMEP does."""
    shape = np.shape(shadow)
    cos_alt = np.cos(altitude * DEG)
    # direct beam: on the cylinder, or per face lit by the sun (face f at azimuth rotation r)
    faces = {'E': (0, azimuth > 360 - t or azimuth <= 180 - t), 'S': (90, 90 - t < azimuth <= 270 - t),
             'W': (180, 180 - t < azimuth <= 360 - t), 'N': (270, azimuth <= 90 - t or azimuth > 270 - t)}
    direct = {}
    if cyl == 1:
        KsideI = shadow * radI * cos_alt
        direct = {f: 0 for f in faces}
    else:
        for f, (rot, lit) in faces.items():
            direct[f] = radI * shadow * cos_alt * np.sin((azimuth - rot + t) * DEG) if lit else 0
        KsideI = shadow * 0
    zero = lambda: np.zeros(shape)
    KsideD, Kside = zero(), zero()
    svf = {'E': (svfE, svfEveg), 'S': (svfS, svfSveg), 'W': (svfW, svfWveg), 'N': (svfN, svfNveg)}
    out = {}
    if anisotropic_diffuse != 1:
        for f, (s_, v_) in svf.items():
            veg, wall = _kvikt(s_, v_)
            w = wall + veg * (1 - psi)
            out[f] = direct[f] + (radD * (1 - w) + albedo * (w * (radG * (1 - F_sh) + radD * F_sh))
                                  + {'E': KupE, 'S': KupS, 'W': KupW, 'N': KupN}[f]) * 0.5
        return out['E'], out['S'], out['W'], out['N'], KsideI, KsideD, Kside
    alt, azi, lum = lv[:, 0], lv[:, 1], lv[:, 2]
    ster = steradians(alt)
    radTot = np.zeros(1)
    for i in range(alt.shape[0]):
        radTot += lum[i] * ster[i] * np.sin(alt[i] * DEG)
    lumChi = (lum * radD) / radTot
    sunlit_surface = ((albedo * (radI * cos_alt) + (radD * 0.5)) / np.pi)
    shaded_surface = ((albedo * radD * 0.5) / np.pi)
    Kref_veg, Kref_sun, Kref_sh = zero(), zero(), zero()
    # face f: diffuse sector (azimuth bounds inclusive above), reflection sector (exclusive), face normal
    diff_sector = {'E': lambda a: a > 360 or a <= 180, 'S': lambda a: 90 < a <= 270,
                   'W': lambda a: 180 < a <= 360, 'N': lambda a: a > 270 or a <= 90}
    refl_sector = {'E': lambda a: a > 360 or a < 180, 'S': lambda a: 90 < a < 270,
                   'W': lambda a: 180 < a < 360, 'N': lambda a: a > 270 or a < 90}
    normal = {'E': 90, 'S': 180, 'W': 270, 'N': 0}
    dRad = {f: zero() for f in faces}
    r_veg = {f: zero() for f in faces}
    r_sun = {f: zero() for f in faces}
    r_sh = {f: zero() for f in faces}
    for k in range(alt.shape[0]):
        ck = np.cos(alt[k] * DEG)
        if cyl == 1:
            KsideD += diffsh[:, :, k] * lumChi[k] * (ck * np.cos(0)) * ster[k]
        else:
            for f in faces:
                if diff_sector[f](azi[k]):
                    dRad[f] += diffsh[:, :, k] * lumChi[k] * (ck * np.cos((normal[f] - azi[k] + t) * DEG)) * ster[k]
        vegsh = (vegshmat[:, :, k] == 0) | (vbshvegshmat[:, :, k] == 0)
        Kref_veg += shaded_surface * vegsh * ster[k] * ck
        wall = ((1 - shmat[:, :, k]) * vbshvegshmat[:, :, k]) == 1
        if cyl == 1:
            sun, shd = shaded_or_sunlit(altitude, azimuth, alt[k], azi[k], asvf)
            Kref_sun += sunlit_surface * sun * wall * ster[k] * ck
            Kref_sh += shaded_surface * shd * wall * ster[k] * ck
            continue
        for f in faces:
            if refl_sector[f](azi[k]):
                r_veg[f] += shaded_surface * ster[k] * ck * vegsh * np.cos((normal[f] - azi[k] + t) * DEG)
        d = np.abs(azimuth - azi[k])
        if 90 < d < 270:
            sun, shd = shaded_or_sunlit(altitude, azimuth, alt[k], azi[k], asvf)
            Kref_sun += sunlit_surface * sun * wall * ster[k] * ck
            Kref_sh += shaded_surface * shd * wall * ster[k] * ck
            for f in faces:
                if refl_sector[f](azi[k]):
                    c = np.cos((normal[f] - azi[k] + t) * DEG)
                    r_sun[f] += sunlit_surface * sun * ster[k] * ck * wall * c
                    r_sh[f] += shaded_surface * shd * ster[k] * ck * wall * c
        else:
            Kref_sh += shaded_surface * wall * ster[k] * ck
            for f in faces:
                if refl_sector[f](azi[k]):
                    r_sh[f] += shaded_surface * ster[k] * ck * wall * np.cos((normal[f] - azi[k] + t) * DEG)
    up = {'E': KupE, 'S': KupS, 'W': KupW, 'N': KupN}
    if cyl == 1:
        Kside = KsideI + KsideD + Kref_sun + Kref_sh + Kref_veg
        return KupE * 0.5, KupS * 0.5, KupW * 0.5, KupN * 0.5, KsideI, KsideD, Kside
    for f in faces:
        out[f] = direct[f] + dRad[f] + r_sun[f] + r_sh[f] + r_veg[f] + up[f] * 0.5
    return out['E'], out['S'], out['W'], out['N'], KsideI, KsideD, Kside


def kup(radI, radD, radG, altitude, svfbuveg, albedo_b, F_sh, gvfalb, gvfalbE, gvfalbS, gvfalbW,
        gvfalbN, gvfalbnosh, gvfalbnoshE, gvfalbnoshS, gvfalbnoshW, gvfalbnoshN):
    """Reflected shortwave from the ground (Kup_veg_2015a, umep_solweig.py:915-927)."""
    s = np.sin(altitude * (np.pi / 180.))
    r = radD * svfbuveg + albedo_b * (1 - svfbuveg) * (radG * (1 - F_sh) + radD * F_sh)
    return tuple(g * radI * s + r * n for g, n in ((gvfalb, gvfalbnosh), (gvfalbE, gvfalbnoshE),
                                                   (gvfalbS, gvfalbnoshS), (gvfalbW, gvfalbnoshW),
                                                   (gvfalbN, gvfalbnoshN)))


def lcyl(esky, sky_patches, Ta, Tgwall, ewall, Lup, shmat, vegshmat, vbshvegshmat, solar_altitude,
         solar_azimuth, asvf):
    """Anisotropic sky longwave on a cylinder (Lcyl_v2022a + model2 + define_patch_characteristics,
    umep_solweig.py:509-541, 946-1029, 1205-1359). Returns (Ldown, Lside, Least, Lwest, Lnorth, Lsouth)."""
    shape = Lup.shape
    alt, azi = sky_patches[:, 0], sky_patches[:, 1]
    skyalt, _ = np.unique(alt, return_counts=True)
    band_emis = 1 - (1 - esky) * np.exp(0.308 * (1.7 - (1 / np.cos((90 - skyalt) * DEG))))
    ster = steradians(alt)
    Ldown_k = np.zeros(alt.shape[0])
    Lside_k = np.zeros(alt.shape[0])
    for a_ in skyalt:
        e = band_emis[skyalt == a_]
        sel = alt == a_
        Ldown_k[sel] = ((e * SBC * ((Ta + 273.15) ** 4)) / np.pi) * ster[sel] * np.sin(a_ * DEG)
        Lside_k[sel] = ((e * SBC * ((Ta + 273.15) ** 4)) / np.pi) * ster[sel] * np.cos(a_ * DEG)
    z = lambda: np.zeros(shape)
    d_sky, d_veg, d_sun, d_sh, d_ref = z(), z(), z(), z(), z()
    s_sky, s_veg, s_sun, s_sh, s_ref = z(), z(), z(), z(), z()
    E, W, N, S = z(), z(), z(), z()

    def cardinal(value, k):
        nonlocal E, W, N, S
        if azi[k] > 360 or azi[k] < 180:
            E = E + value * np.cos((90 - azi[k]) * DEG)
        if 90 < azi[k] < 270:
            S = S + value * np.cos((180 - azi[k]) * DEG)
        if 180 < azi[k] < 360:
            W = W + value * np.cos((270 - azi[k]) * DEG)
        if azi[k] > 270 or azi[k] < 90:
            N = N + value * np.cos((0 - azi[k]) * DEG)

    veg_surface = ((ewall * SBC * ((Ta + 273.15) ** 4)) / np.pi)
    sunlit_surface = ((ewall * SBC * ((Ta + Tgwall + 273.15) ** 4)) / np.pi)
    shaded_surface = ((ewall * SBC * ((Ta + 273.15) ** 4)) / np.pi)
    for k in range(alt.shape[0]):
        sky = (shmat[:, :, k] == 1) & (vegshmat[:, :, k] == 1)
        d_sky += sky * Ldown_k[k]
        s_sky += sky * Lside_k[k]
        veg = (vegshmat[:, :, k] == 0) | (vbshvegshmat[:, :, k] == 0)
        s_veg += veg_surface * ster[k] * np.cos(alt[k] * DEG) * veg
        d_veg += veg_surface * ster[k] * np.sin(alt[k] * DEG) * veg
        cardinal(sky * Lside_k[k], k)
        cardinal(veg_surface * ster[k] * np.cos(alt[k] * DEG) * veg, k)
        wall = ((1 - shmat[:, :, k]) * vbshvegshmat[:, :, k]) == 1
        diff = np.abs(solar_azimuth - azi[k])
        if (diff > 90) and (diff < 270) and (solar_altitude > 0):
            sun, shd = shaded_or_sunlit(solar_altitude, solar_azimuth, alt[k], azi[k], asvf)
            s_sun += sunlit_surface * sun * ster[k] * np.cos(alt[k] * DEG) * wall
            s_sh += shaded_surface * shd * ster[k] * np.cos(alt[k] * DEG) * wall
            d_sun += sunlit_surface * sun * ster[k] * np.sin(alt[k] * DEG) * wall
            d_sh += shaded_surface * shd * ster[k] * np.sin(alt[k] * DEG) * wall
            cardinal(sunlit_surface * sun * ster[k] * np.cos(alt[k] * DEG) * wall, k)
            cardinal(shaded_surface * shd * ster[k] * np.cos(alt[k] * DEG) * wall, k)
        else:
            s_sh += shaded_surface * ster[k] * np.cos(alt[k] * DEG) * wall
            d_sh += shaded_surface * ster[k] * np.sin(alt[k] * DEG) * wall
            cardinal(shaded_surface * ster[k] * np.cos(alt[k] * DEG) * wall, k)
    refl = (((d_sky + Lup) * (1 - ewall) * 0.5) / np.pi)
    for k in range(alt.shape[0]):
        obst = (shmat[:, :, k] == 0) | (vegshmat[:, :, k] == 0) | (vbshvegshmat[:, :, k] == 0)
        s_ref += refl * ster[k] * np.cos(alt[k] * DEG) * obst
        d_ref += refl * ster[k] * np.sin(alt[k] * DEG) * obst
        cardinal(refl * ster[k] * np.cos(alt[k] * DEG) * obst, k)
    Lside = s_sky + s_veg + s_sh + s_sun + s_ref
    Ldown = d_sky + d_veg + d_sh + d_sun + d_ref
    return Ldown, Lside, E, W, N, S


# --------------------------------------------------------------------------- the step

def step(p, hour, state):
    """One SOLWEIG 2022a step (umep_solweig.py:50-372) for the configuration series.py runs.

    ``p``: static dict (dsm, vegdem, vegdem2, bush, walls, dirwalls, buildings, svf planes, asvf,
    svfalfa, svfbuveg, diffsh/shmat/vegshmat/vbshvegshmat [rows, cols, 153], land-cover planes and
    scalars). ``hour``: altitude, azimuth, zen, jday, dectime, altmax, Ta, RH, radG, P, psi,
    Twater. ``state``: CI, firstdaytime, timeadd, timestepdec, Tgmap1(E,S,W,N), TgOut1.
    Returns (outputs dict, new state dict).
    """
    rows, cols = p['dsm'].shape
    Ta, RH, radG, P = hour['Ta'], hour['RH'], hour['radG'], hour['P']
    altitude, azimuth, zen, jday = hour['altitude'], hour['azimuth'], hour['zen'], hour['jday']
    dectime, altmax, psi = hour['dectime'], hour['altmax'], hour['psi']
    CI = state['CI']
    firstdaytime, timeadd, timestepdec = state['firstdaytime'], state['timeadd'], state['timestepdec']
    Tgmap = {k: state[k] for k in ('Tgmap1', 'Tgmap1E', 'Tgmap1S', 'Tgmap1W', 'Tgmap1N', 'TgOut1')}
    albedo_b, ewall = p['albedo_b'], p['ewall']
    absK, absL, Fside, Fup, Fcyl = p['absK'], p['absL'], p['Fside'], p['Fup'], p['Fcyl']
    emis_grid, alb_grid = p['emis_grid'], p['alb_grid']
    landcover = p['landcover']
    _, _, _, SNUP = daylen(jday, p['latitude'])
    ea = 6.107 * 10 ** ((7.5 * Ta) / (237.3 + Ta)) * (RH / 100.)
    msteg = 46.5 * (ea / (Ta + 273.15))
    esky = (1 - (1 + msteg) * np.exp(-((1.2 + 3.0 * msteg) ** 0.5))) + p.get('elvis', 0)
    out = {}
    lv = None
    if altitude > 0:
        I0, CI, Kt, _, _ = clearness_index(zen, jday, Ta, RH / 100., radG, p['latitude'], P)
        if (CI > 1) or (CI == np.inf):
            CI = 1
        radI, radD = diffuse_fraction(radG, altitude, Kt, Ta, RH)
        lv, _, _ = perez(zen * (180 / np.pi), azimuth, radD, radI, jday, 2, clip=p.get('perez_clip', False))
        ani = np.zeros((rows, cols))
        for k in range(lv.shape[0]):
            ani += p['diffsh'][:, :, k] * lv[k, 2]
        dRad = ani * radD
        vegsh, sh, vbshvegsh, wallsh, wallsun, wallshve, facesh, facesun = shadows(
            p['dsm'], p['vegdem'], p['vegdem2'], azimuth, altitude, p['scale'], p['amaxvalue'],
            p['bush'], p['walls'], p['dirwalls'] * np.pi / 180.)
        shadow = sh - (1 - vegsh) * (1 - psi)
        Tgamp = p['TgK'] * altmax + p['Tstart']
        Tgampwall = p['TgK_wall'] * altmax + p['Tstart_wall']
        frac = dectime - np.floor(dectime)
        Tg = Tgamp * np.sin(((frac - SNUP / 24) / (p['TmaxLST'] / 24 - SNUP / 24)) * np.pi / 2)
        Tgwall = Tgampwall * np.sin(((frac - SNUP / 24) / (p['TmaxLST_wall'] / 24 - SNUP / 24)) * np.pi / 2)
        if Tgwall < 0:
            Tgwall = 0
        radI0, radD0 = diffuse_fraction(I0, altitude, 1., Ta, RH)
        corr = 0.1473 * np.log(90 - (zen / np.pi * 180)) + 0.3454
        CI_Tg = (radG / radI0) + (1 - corr)
        if (CI_Tg > 1) or (CI_Tg == np.inf):
            CI_Tg = 1
        radG0 = radI0 * (np.sin(altitude * DEG)) + radD0
        CI_TgG = (radG / radG0) + (1 - corr)
        if (CI_TgG > 1) or (CI_TgG == np.inf):
            CI_TgG = 1
        Tg = Tg * CI_TgG
        Tgwall = Tgwall * CI_TgG
        if landcover == 1:
            Tg = np.array(Tg, dtype=np.float64)
            Tg[Tg < 0] = 0
        Tg = np.array(np.broadcast_to(Tg, (rows, cols)), dtype=np.float64)
        g = gvf(wallsun, p['walls'], p['buildings'], p['scale'], shadow, p['first'], p['second'],
                p['dirwalls'], Tg, Tgwall, Ta, emis_grid, ewall, alb_grid, albedo_b, p.get('Twater'),
                p['lc_grid'], landcover)
        (gvfLup, gvfalb, gvfalbnosh, gvfLupE, gvfalbE, gvfalbnoshE, gvfLupS, gvfalbS, gvfalbnoshS,
         gvfLupW, gvfalbW, gvfalbnoshW, gvfLupN, gvfalbN, gvfalbnoshN, gvfSum, gvfNorm) = g
        Lup, _, Tgmap['Tgmap1'] = wave_delay(gvfLup, firstdaytime, timeadd, timestepdec, Tgmap['Tgmap1'])
        LupE, _, Tgmap['Tgmap1E'] = wave_delay(gvfLupE, firstdaytime, timeadd, timestepdec, Tgmap['Tgmap1E'])
        LupS, _, Tgmap['Tgmap1S'] = wave_delay(gvfLupS, firstdaytime, timeadd, timestepdec, Tgmap['Tgmap1S'])
        LupW, _, Tgmap['Tgmap1W'] = wave_delay(gvfLupW, firstdaytime, timeadd, timestepdec, Tgmap['Tgmap1W'])
        LupN, _, Tgmap['Tgmap1N'] = wave_delay(gvfLupN, firstdaytime, timeadd, timestepdec, Tgmap['Tgmap1N'])
        TgTemp = Tg * shadow + Ta
        TgOut, timeadd, Tgmap['TgOut1'] = wave_delay(TgTemp, firstdaytime, timeadd, timestepdec, Tgmap['TgOut1'])
        F_sh = cylindric_wedge(zen, p['svfalfa'])
        F_sh[np.isnan(F_sh)] = 0.5
        Kdown = radI * shadow * np.sin(altitude * (np.pi / 180)) + dRad + albedo_b * (1 - p['svfbuveg']) * \
            (radG * (1 - F_sh) + radD * F_sh)
        Kup_, KupE, KupS, KupW, KupN = kup(radI, radD, radG, altitude, p['svfbuveg'], albedo_b, F_sh,
                                          gvfalb, gvfalbE, gvfalbS, gvfalbW, gvfalbN, gvfalbnosh,
                                          gvfalbnoshE, gvfalbnoshS, gvfalbnoshW, gvfalbnoshN)
        Keast, Ksouth, Kwest, Knorth, KsideI, KsideD, Kside = kside_cylinder(
            radI, radD, shadow, azimuth, altitude, albedo_b, KupE, KupS, KupW, KupN, lv,
            p['diffsh'], p['asvf'], p['shmat'], p['vegshmat'], p['vbshvegshmat'])
        firstdaytime = 0
        out.update(shadow=shadow, sh=sh, vegsh=vegsh, vbshvegsh=vbshvegsh, wallsun=wallsun,
                   wallsh=wallsh, wallshve=wallshve, facesh=facesh, facesun=facesun,
                   gvfLup=gvfLup, gvfalb=gvfalb, gvfalbnosh=gvfalbnosh, gvfNorm=gvfNorm, F_sh=F_sh,
                   radI=radI, radD=radD, CI_Tg=CI_Tg, CI_TgG=CI_TgG, Tgwall=Tgwall, I0=I0)
    else:
        Tgwall = 0
        z = np.zeros((rows, cols))
        Kdown = Kup_ = Keast = Ksouth = Kwest = Knorth = KsideI = KsideD = Kside = dRad = z
        Tg = np.zeros((rows, cols))
        shadow = np.zeros((rows, cols))
        Lup = SBC * emis_grid * ((z + Ta + Tg + 273.15) ** 4)
        if landcover == 1:
            Lup = np.array(np.broadcast_to(Lup, (rows, cols)), dtype=np.float64)
            Lup[p['lc_grid'] == 3] = SBC * 0.98 * (p.get('Twater') + 273.15) ** 4
        LupE = LupS = LupW = LupN = Lup
        TgOut = Ta + Tg
        timeadd = 0
        firstdaytime = 1
        out.update(shadow=shadow, radI=hour.get('radI', 0.0), radD=hour.get('radD', 0.0),
                   Tgwall=0.0, I0=0)
    if lv is None:
        alt_p, azi_p = create_patches(2)
        patches = np.column_stack((alt_p, azi_p, np.zeros_like(alt_p)))
    else:
        patches = lv.copy()
    if CI < 0.95:
        esky = CI * esky + (1 - CI) * 1.
    Ldown, Lside, Least_, Lwest_, Lnorth_, Lsouth_ = lcyl(
        esky, patches, Ta, Tgwall, ewall, Lup, p['shmat'], p['vegshmat'], p['vbshvegshmat'],
        altitude, azimuth, p['asvf'])
    Least, Lsouth, Lwest, Lnorth = LupE * 0.5, LupS * 0.5, LupW * 0.5, LupN * 0.5
    Sstr = absK * (Kside * Fcyl + (Kdown + Kup_) * Fup + (Knorth + Keast + Ksouth + Kwest) * Fside) + absL * \
        ((Ldown + Lup) * Fup + Lside * Fcyl + (Lnorth + Least + Lsouth + Lwest) * Fside)
    Tmrt = np.sqrt(np.sqrt((Sstr / (absL * SBC)))) - 273.2
    out.update(Tmrt=Tmrt, Kdown=Kdown, Kup=Kup_, Ldown=Ldown, Lup=Lup, Tg=Tg, Lside=Lside,
               Kside=Kside, KsideI=KsideI, KsideD=KsideD, dRad=dRad, esky=esky, ea=ea, CI=CI,
               Keast=Keast, Ksouth=Ksouth, Kwest=Kwest, Knorth=Knorth,
               Least=Least + Least_, Lsouth=Lsouth + Lsouth_, Lwest=Lwest + Lwest_,
               Lnorth=Lnorth + Lnorth_, LupE=LupE, LupS=LupS, LupW=LupW, LupN=LupN,
               TgOut=TgOut, Sstr=Sstr)
    new_state = dict(CI=CI, firstdaytime=firstdaytime, timeadd=timeadd, timestepdec=timestepdec, **Tgmap)
    return out, new_state


# ----------------------------------------------------------------------------- per-step sub-models

def lside_weights(svf, svfveg, svfaveg, total=4.4897):
    """View-factor weights of wall, vegetation, sky and reflection for one cardinal direction
    (Lvikt_veg, umep_solweig.py:1183-1198): (veg, wall, sky, refl)."""
    def p(x):
        return 63.227 * x ** 6 - 161.51 * x ** 5 + 156.91 * x ** 4 - 70.424 * x ** 3 + 16.773 * x ** 2 - 0.4863 * x
    wall = (total - p(svf)) / total - (total - p(svfaveg)) / total
    both = svfveg + svf - 1
    sky = p(both) / total
    refl = (total - p(both)) / total
    veg = refl - wall
    return veg, wall, sky, refl


def lside(svfS, svfW, svfN, svfE, svfEveg, svfSveg, svfWveg, svfNveg, svfEaveg, svfSaveg, svfWaveg,
          svfNaveg, azimuth, altitude, Ta, Tw, SBC, ewall, Ldown, esky, t, F_sh, CI, LupE, LupS, LupW,
          LupN, anisotropic_longwave):
    """Longwave on the four cardinal faces of a standing body (Lside_veg_v2022a,
    umep_solweig.py:1032-1180): returns (Least, Lsouth, Lwest, Lnorth)."""
    Fs = 2 * F_sh - 1
    c = 1 - CI
    T4 = (Ta + 273.15) ** 4
    allsky = esky * SBC * T4 * (1 - c) + c * SBC * T4
    faces = (  # (svf, veg, aveg, Lup, lit(azimuth), azimuth offset of the face)
        (svfE, svfEveg, svfEaveg, LupE, (azimuth > (180 - t)) and (azimuth <= (360 - t)), azimuth - 180 + t),
        (svfS, svfSveg, svfSaveg, LupS, (azimuth <= (90 - t)) or (azimuth > (270 - t)), azimuth - 270 + t),
        (svfW, svfWveg, svfWaveg, LupW, (azimuth > (360 - t)) or (azimuth <= (180 - t)), azimuth + t),
        (svfN, svfNveg, svfNaveg, LupN, (azimuth > (90 - t)) and (azimuth <= (270 - t)), azimuth - 90 + t))
    out = []
    for svf, veg, aveg, up, lit, azi in faces:
        if anisotropic_longwave == 1:
            out.append(up * 0.5)
            continue
        wveg, wwall, wsky, wrefl = lside_weights(svf, veg, aveg)
        if altitude > 0 and lit:
            alfa = np.arcsin(np.exp(np.log(1 - svf) / 2))
            a_b = np.arctan(alfa)
            b_b = np.arctan(np.tan(alfa * Fs))
            betasun = (a_b - b_b) / 2 + b_b
            sun = SBC * ewall * ((Ta + 273.15 + Tw * np.sin(azi * (np.pi / 180))) ** 4) * wwall * (1 - Fs) * \
                np.cos(betasun) * 0.5
            shade = SBC * ewall * T4 * wwall * Fs * 0.5
        else:
            sun = 0
            shade = SBC * ewall * T4 * wwall * 0.5
        sky = ((svf + veg - 1) * allsky) * wsky * 0.5
        vegl = SBC * ewall * T4 * wveg * 0.5
        refl = (Ldown + up) * wrefl * (1 - ewall) * 0.5
        out.append(sky + sun + shade + vegl + up * 0.5 + refl)
    return tuple(out)


def surface_temperature(TgK, Tstart, TgK_wall, Tstart_wall, TmaxLST, TmaxLST_wall, altmax, dectime, SNUP,
                        CI_TgG, landcover=0):
    """Ground and wall temperature wave (umep_solweig.py:144-170, as in step): (Tg, Tgwall)."""
    frac = dectime - np.floor(dectime)
    Tg = (TgK * altmax + Tstart) * np.sin(((frac - SNUP / 24) / (TmaxLST / 24 - SNUP / 24)) * np.pi / 2)
    Tgw = (TgK_wall * altmax + Tstart_wall) * np.sin(((frac - SNUP / 24) / (TmaxLST_wall / 24 - SNUP / 24)) *
                                                    np.pi / 2)
    Tgw = np.where(Tgw < 0, 0.0, Tgw)
    Tg = Tg * CI_TgG
    Tgw = Tgw * CI_TgG
    if landcover == 1:
        Tg = np.where(Tg < 0, 0.0, Tg)
    return Tg, Tgw


def assemble_tmrt(Kdown, Kup, Knorth, Keast, Ksouth, Kwest, Ldown, Lup, Lnorth, Least, Lsouth, Lwest, KsideI,
                  Kside, Lside, absK, absL, Fside, Fup, Fcyl, cyl, anisotropic_sky, SBC=5.67051e-8):
    """Mean radiant temperature from the absorbed flux (umep_solweig.py:345-360): (Tmrt, Sstr)."""
    K4 = Knorth + Keast + Ksouth + Kwest
    L4 = Lnorth + Least + Lsouth + Lwest
    if cyl == 1 and anisotropic_sky == 1:
        S = absK * (Kside * Fcyl + (Kdown + Kup) * Fup + K4 * Fside) + \
            absL * ((Ldown + Lup) * Fup + Lside * Fcyl + L4 * Fside)
    elif cyl == 1:
        S = absK * (KsideI * Fcyl + (Kdown + Kup) * Fup + K4 * Fside) + absL * ((Ldown + Lup) * Fup + L4 * Fside)
    else:
        S = absK * ((Kdown + Kup) * Fup + K4 * Fside) + absL * ((Ldown + Lup) * Fup + L4 * Fside)
    return np.sqrt(np.sqrt(S / (absL * SBC))) - 273.2, S
