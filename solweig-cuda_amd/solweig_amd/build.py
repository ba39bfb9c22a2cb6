"""Build libsolweig_hip.so in place with hipcc for gfx950.

    python -m solweig_amd.build          (from solweig-cuda_amd/)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(HERE, 'csrc', 'solweig_hip.hip')]
OUTPUT = os.path.join(HERE, 'libsolweig_hip.so')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         # packed FP32 VALU forms stay disabled (they have returned wrong values on these parts)
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         # hardware float atomics (the adjoint's accumulations into device-memory buffers)
         '-munsafe-fp-atomics',
         # no machine-level loop-invariant code motion: in the day pixel pass (one loop over the
         # hours around the whole per-hour body) it hoists immediates into vector registers held
         # across the loop (158 instead of 105 VGPRs); the other kernels are unaffected in time
         '-mllvm', '-disable-machine-licm']


def build(verbose=False):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    tmp = OUTPUT + '.tmp'
    cmd = [hipcc, *FLAGS, '-o', tmp, *SOURCES]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f'hipcc failed ({res.returncode})')
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == '__main__':
    print(build(verbose=True))
