"""Writes spa_terms.inc: the periodic-term tables of the NREL Solar Position Algorithm (Reda &
Andreas 2003, tables A4.2 and A4.3, as spa_tables.py holds them) as device constant arrays for
the solar-position kernels.

    python gen_spa_terms.py          (from this directory)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from solweig_amd import spa_tables as T  # noqa: E402

GROUPS = ('L0', 'L1', 'L2', 'L3', 'L4', 'L5', 'B0', 'B1', 'R0', 'R1', 'R2', 'R3', 'R4')


def render():
    """The text of spa_terms.inc."""
    rows, spans = [], []
    for name in GROUPS:
        g = getattr(T, name)
        spans.append((len(rows), len(g)))
        rows.extend(g)
    lines = ['/* generated by gen_spa_terms.py from spa_tables.py: the L0-L5, B0-B1, R0-R4 rows (A, B, C) of',
             ' * A cos(B + C JME) back to back, each group\'s (first row, row count), and the nutation',
             ' * argument multipliers and coefficients (a, b, c, d). */',
             f'constexpr int SPA_ROWS = {len(rows)}, SPA_NUT_ROWS = {len(T.NUTATION_ARGS)};',
             '__constant__ double SPA_TERMS[SPA_ROWS][3] = {']
    lines += ['    {%r, %r, %r},' % tuple(float(v) for v in r) for r in rows]
    lines += ['};', '__constant__ int SPA_GROUP[13][2] = {' + ', '.join('{%d, %d}' % s for s in spans) + '};',
              '__constant__ double SPA_NUT_ARGS[SPA_NUT_ROWS][5] = {']
    lines += ['    {%r, %r, %r, %r, %r},' % tuple(float(v) for v in r) for r in T.NUTATION_ARGS]
    lines += ['};', '__constant__ double SPA_NUT_COEFFS[SPA_NUT_ROWS][4] = {']
    lines += ['    {%r, %r, %r, %r},' % tuple(float(v) for v in r) for r in T.NUTATION_COEFFS]
    lines += ['};']
    return '\n'.join(lines) + '\n'


def main():
    with open(os.path.join(HERE, 'spa_terms.inc'), 'w') as fh:
        fh.write(render())


if __name__ == '__main__':
    main()
