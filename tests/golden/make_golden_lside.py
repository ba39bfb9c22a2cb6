"""Golden vectors of the isotropic cardinal longwave (Lside_veg_v2022a with
anisotropic_longwave = 0), the one per-step sub-model no closed form pins.

Run in the build container only:

    python tests/golden/make_golden_lside.py /root/reference

Inputs are seeded random planes; only inputs and outputs are stored.
"""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(reference):
    spec = importlib.util.spec_from_file_location('umep_solweig', os.path.join(reference, 'src', 'umep_solweig.py'))
    U = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(U)
    rng = np.random.default_rng(71)
    shape = (9, 11)
    out = {}
    names = ('svfS', 'svfW', 'svfN', 'svfE', 'svfEveg', 'svfSveg', 'svfWveg', 'svfNveg', 'svfEaveg', 'svfSaveg',
             'svfWaveg', 'svfNaveg')
    for case, (azimuth, altitude, CI) in enumerate(((135.0, 40.0, 0.8), (250.0, 12.0, 1.0), (20.0, -5.0, 0.6),
                                                    (315.0, 55.0, 0.97))):
        planes = {n: rng.uniform(0.2, 0.95, size=shape) for n in names}
        Tw = rng.uniform(0.0, 8.0, size=shape)
        Ldown = rng.uniform(300, 420, size=shape)
        F_sh = rng.uniform(0.0, 1.0, size=shape)
        Lup = {k: rng.uniform(380, 520, size=shape) for k in 'ESWN'}
        args = [planes[n] for n in names] + [azimuth, altitude, 24.5, Tw, 5.67051e-8, 0.9, Ldown, 0.83, 0.0, F_sh, CI,
                                               Lup['E'], Lup['S'], Lup['W'], Lup['N'], 0]
        res = U.Lside_veg_v2022a(*args)
        p = f'c{case}_'
        for n in names:
            out[p + n] = planes[n]
        out[p + 'scalars'] = np.array([azimuth, altitude, 24.5, 5.67051e-8, 0.9, 0.83, 0.0, CI])
        out[p + 'Tw'], out[p + 'Ldown'], out[p + 'F_sh'] = Tw, Ldown, F_sh
        for k in 'ESWN':
            out[p + 'Lup' + k] = Lup[k]
        out[p + 'out'] = np.stack([np.asarray(r, dtype=np.float64) for r in res])
    np.savez_compressed(os.path.join(HERE, 'umep_lside_golden.npz'), **out)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
