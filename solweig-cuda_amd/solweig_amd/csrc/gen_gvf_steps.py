"""Writes gvf_steps.inc: the (row, column) shifts of the 18 ground-view search directions for
the first 22 steps (sky.gvf_geometry: sunonsurface_2018a, umep_solweig.py:1477-1492, azimuth
5 + 20 j degrees). Every scale's table starts with these steps; scale 1 walks exactly 22.

    python gen_gvf_steps.py          (from this directory)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from solweig_amd.sky import gvf_geometry  # noqa: E402

STEPS = 22


def main():
    o = gvf_geometry(1.0)[0]
    assert o.shape[1] >= STEPS
    lines = ['/* generated by gen_gvf_steps.py: GV_STEPS[direction][step] = {row, column} shift */']
    for j in range(18):
        lines.append('{' + ', '.join('{%d, %d}' % (o[j, n, 0], o[j, n, 1]) for n in range(STEPS)) + '},')
    with open(os.path.join(HERE, 'gvf_steps.inc'), 'w') as fh:
        fh.write('\n'.join(lines) + '\n')


if __name__ == '__main__':
    main()
