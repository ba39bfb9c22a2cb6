"""CPU oracle for the SOLWEIG 2022a per-pixel-hour radiation path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import this package, and only as the checker or the timed CPU baseline.
The product (``solweig-cuda_amd/solweig_amd``) never imports it and has no CPU fallback.

Contents (each function cites the reference lines it restates):

* ``spa``    - NREL solar position and the met-row sun sampling (umep_solweig.py:2412-3521).
* ``umep``   - the 2022a step and its sub-models in FP64 NumPy, following umep_solweig.py
               (cyl=1, anisotropic sky, vegetation, global-only branches, as series.py runs it).
* ``svf``    - the 153-patch sky-view visibility masks and band reduction of the reference's
               svf.py:19-114 / geometry.cu:77-159.
* ``brute``  - a per-pixel loop form of the shadow march, used only to pin ``umep.shadows``.
* ``utci``   - the UTCI polynomial in FP64 (utci.cu:1-241) and its per-hour factored form
               (utci.cu:243-523).
* ``day``    - static preparation (series.py:252-349) and the day loop (series.py:396-495).

Parity pins (see tests/): golden vectors written by importing the reference's umep_solweig.py in
the build container (tests/golden/make_golden.py), closed forms (flat-ground SVF = 1, box shadow
length, uniform-field GVF identity, open-sky Ldown, lag recursion), and the brute-force march.
``utci`` is pinned by golden vectors evaluated from the reference's own expression and by a
published UTCI example value.
"""
