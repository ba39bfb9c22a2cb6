"""Rebuild oracle static inputs and sun tables from a golden day fixture."""
import numpy as np

from oracle import day as oday

SVF_KEYS = ('svf', 'svfaveg', 'svfE', 'svfEaveg', 'svfEveg', 'svfN', 'svfNaveg', 'svfNveg',
            'svfS', 'svfSaveg', 'svfSveg', 'svfveg', 'svfW', 'svfWaveg', 'svfWveg')


def unpack(g, key, patches=153):
    return np.unpackbits(g[key], axis=-1)[..., :patches].astype(np.float32)


def static_from_golden(g, prefix):
    shm, vgm, vbm = (unpack(g, prefix + k) for k in ('shmat', 'vegmat', 'vbmat'))
    svf15 = g[prefix + 'svf15'].astype(np.float64)
    svf = tuple(svf15) + (vgm, vbm, shm, None)
    lc = g[prefix + 'landcover'] if prefix + 'landcover' in g else None
    alb = g[prefix + 'alb_grid'] if prefix + 'alb_grid' in g else None
    emis = g[prefix + 'emis_grid'] if prefix + 'emis_grid' in g else None
    loc = g[prefix + 'loc']
    p = oday.prepare(g[prefix + 'dsm'], g[prefix + 'trees'], g[prefix + 'dem'], g[prefix + 'walls'],
                     g[prefix + 'aspect'], float(loc[0]), landcover=lc, alb_grid=alb, emis_grid=emis,
                     svf=svf)
    return p


def sun_from_golden(g, prefix):
    s = g[prefix + 'sun']
    return dict(altitude=s[0], azimuth=s[1], zen=s[2], jday=s[3], dectime=s[4], altmax=s[5])
