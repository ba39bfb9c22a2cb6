"""Golden vectors of the anisotropic-sky sub-models over random visibility masks:
Kside_veg_v2022a in all four (cyl, anisotropic_diffuse) branches and Lcyl_v2022a by day and by
night. The step-level vectors only exercise cyl = 1 with anisotropic diffuse and the series'
masks; these pin the box and isotropic branches and arbitrary mask patterns.

Run in the build container only:

    python tests/golden/make_golden_aniso.py /root/reference

Inputs are seeded random planes and masks; only inputs and outputs are stored.
"""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPE = (7, 9)


def _masks(rng, patches):
    shape = (*SHAPE, patches)
    sh = (rng.random(shape) < 0.6).astype(np.float64)
    veg = (rng.random(shape) < 0.7).astype(np.float64)
    vb = (rng.random(shape) < 0.75).astype(np.float64)
    diffsh = np.where(rng.random(shape) < 0.5, sh, rng.uniform(0.0, 1.0, shape))
    return sh, veg, vb, diffsh


def main(reference):
    spec = importlib.util.spec_from_file_location('umep_solweig', os.path.join(reference, 'src', 'umep_solweig.py'))
    U = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(U)
    rng = np.random.default_rng(2024)
    out = {}
    rows, cols = SHAPE
    # (zen deg, azimuth, altitude, radI, radD, radG, t, psi, albedo, cyl, anisotropic)
    kcases = [(50.0, 135.0, 40.0, 610.0, 140.0, 600.0, 0.0, 0.97, 0.2, 1, 1),
              (70.0, 250.0, 20.0, 380.0, 110.0, 240.0, 0.0, 0.5, 0.25, 0, 1),
              (35.0, 200.0, 55.0, 700.0, 90.0, 660.0, 15.0, 0.8, 0.18, 0, 1),
              (60.0, 95.0, 30.0, 450.0, 160.0, 390.0, 0.0, 0.97, 0.2, 1, 0),
              (40.0, 300.0, 50.0, 520.0, 120.0, 510.0, 10.0, 0.6, 0.3, 0, 0)]
    for c, (zen, az, alt, radI, radD, radG, t, psi, albedo, cyl, ani) in enumerate(kcases):
        lv, _, _ = U.Perez_v3(zen, az, radD, radI, 172, 1, 2)
        lv = np.asarray(lv, dtype=np.float64)
        sh, veg, vb, diffsh = _masks(rng, lv.shape[0])
        planes = {n: rng.uniform(0.2, 0.95, SHAPE) for n in ('svfS', 'svfW', 'svfN', 'svfE', 'svfEveg', 'svfSveg',
                                                           'svfWveg', 'svfNveg')}
        shadow = np.where(rng.random(SHAPE) < 0.5, 1.0, rng.uniform(0.0, 1.0, SHAPE))
        F_sh = rng.uniform(0.0, 1.0, SHAPE)
        kup = {k: rng.uniform(20.0, 140.0, SHAPE) for k in 'ESWN'}
        asvf = rng.uniform(0.2, 1.5, SHAPE)
        res = U.Kside_veg_v2022a(radI, radD, radG, shadow, planes['svfS'], planes['svfW'], planes['svfN'],
                                 planes['svfE'], planes['svfEveg'], planes['svfSveg'], planes['svfWveg'],
                                 planes['svfNveg'], az, alt, psi, t, albedo, F_sh, kup['E'], kup['S'], kup['W'],
                                 kup['N'], cyl, lv, ani, diffsh, rows, cols, asvf, sh, veg, vb)
        p = f'k{c}_'
        out[p + 'scalars'] = np.array([radI, radD, radG, az, alt, psi, t, albedo, cyl, ani])
        out[p + 'lv'] = lv
        out[p + 'masks'] = np.stack([sh, veg, vb]).astype(np.uint8)
        out[p + 'diffsh'] = diffsh
        for n, v in planes.items():
            out[p + n] = v
        out[p + 'shadow'], out[p + 'F_sh'], out[p + 'asvf'] = shadow, F_sh, asvf
        for k in 'ESWN':
            out[p + 'Kup' + k] = kup[k]
        out[p + 'out'] = np.stack([np.broadcast_to(np.asarray(r, dtype=np.float64), SHAPE) for r in res])
    # (esky, Ta, Tgwall, ewall, solar altitude, solar azimuth); day patches from Perez, night the vault
    lcases = [(0.82, 24.0, 6.5, 0.9, 35.0, 160.0), (0.75, 12.0, 0.0, 0.92, -8.0, 20.0),
              (0.88, 30.0, 11.0, 0.9, 62.0, 250.0)]
    for c, (esky, Ta, Tgwall, ewall, salt, sazi) in enumerate(lcases):
        lv, _, _ = U.Perez_v3(90.0 - max(salt, 1.0), sazi, 100.0, 400.0, 172, 1, 2)
        patches = np.asarray(lv, dtype=np.float64)
        sh, veg, vb, _ = _masks(rng, patches.shape[0])
        Lup = rng.uniform(380.0, 520.0, SHAPE)
        asvf = rng.uniform(0.2, 1.5, SHAPE)
        res = U.Lcyl_v2022a(esky, patches, Ta, Tgwall, ewall, Lup, sh, veg, vb, salt, sazi, rows, cols, asvf)
        p = f'l{c}_'
        out[p + 'scalars'] = np.array([esky, Ta, Tgwall, ewall, salt, sazi])
        out[p + 'patches'] = patches
        out[p + 'masks'] = np.stack([sh, veg, vb]).astype(np.uint8)
        out[p + 'Lup'], out[p + 'asvf'] = Lup, asvf
        out[p + 'out'] = np.stack([np.asarray(r, dtype=np.float64) for r in res])
    np.savez_compressed(os.path.join(HERE, 'umep_aniso_golden.npz'), **out)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
