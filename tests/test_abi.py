"""C ABI surface: the library loads, exports every declared symbol, and the Python mirrors of
the header's enums and struct match it (no GPU needed)."""
import ctypes
import os
import re

import pytest

from solweig_amd import runtime as rt

HEADER = os.path.join(os.path.dirname(__file__), '..', 'include', 'solweig_hip.h')


def header_text():
    return open(HEADER).read()


def test_library_exports_every_declared_symbol():
    if not os.path.exists(rt.LIBRARY):
        pytest.skip('library not built')
    L = ctypes.CDLL(rt.LIBRARY)
    declared = re.findall(r'^\s*(?:int|const char \*)\s*(sw_\w+)\(', header_text(), flags=re.M)
    assert declared, 'no declarations parsed'
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(rt.EXPORTED)
    assert rt.lib().sw_version() >= 100


def _enum_values(block_start):
    text = header_text()
    i = text.rfind('enum {', 0, text.index(block_start))
    body = text[i:text.index('};', i)]
    body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
    vals, cur = {}, -1
    for tok in [t.strip() for t in body[body.index('{') + 1:].split(',') if t.strip()]:
        if '=' in tok:
            name, expr = (x.strip() for x in tok.split('='))
            cur = eval(expr, {}, {k.replace('SWH_', ''): v for k, v in vals.items()} | vals)
        else:
            name, cur = tok, cur + 1
        vals[name] = cur
    return vals


def test_hour_record_layout_matches_header():
    vals = _enum_values('SWH_DAY = 0')
    for name, idx in rt.H.items():
        assert vals['SWH_' + name] == idx, name
    assert vals['SW_HOUR_STRIDE'] == rt.SW_HOUR_STRIDE


def test_domain_struct_is_336_bytes():
    assert ctypes.sizeof(rt.SwDomain) == 336
    text = header_text()[header_text().index('typedef struct sw_domain'):header_text().index('} sw_domain;')]
    fields = re.findall(r'\*(\w+)', text)
    names = [f[0] for f in rt.SwDomain._fields_ if f[0] not in ('rows', 'cols', 'pixels', 'landcover',
                                                                'dsm_raise', 'tile_cols', 'hmax_cols')]
    assert fields == names
    assert [f[0] for f in rt.SwDomain._fields_][-8:] == ['hmax_cols', 'gvf_len', 'gvf_nosh', 'march_trees',
                                                         'ptab_s', 'gvf_edge', 'diffuse_tab', 'diffuse']


def test_adjoint_slots_and_struct_match_header():
    vals = _enum_values('SWA_TA = 0')
    for name, idx in rt.ADJ.items():
        assert vals['SWA_' + name] == idx, name
    assert vals['SWA_SLOTS'] == rt.ADJ_SLOTS
    assert vals['SWA_PARTIAL_SLOTS'] == rt.ADJ_PARTIAL_SLOTS
    pv = _enum_values('SWA_PATCH_SKY_KDOWN = 0')
    for name, idx in rt.ADJ_PATCH.items():
        assert pv['SWA_PATCH_' + name] == idx
    assert ctypes.sizeof(rt.SwAdjoint) == 128
    text = header_text()
    block = text[text.index('typedef struct sw_adjoint'):text.index('} sw_adjoint;')]
    fields = re.findall(r'\*(\w+);', block)
    assert fields == [f[0] for f in rt.SwAdjoint._fields_][:-1]



@pytest.mark.parametrize('scale', [1.0, 0.5, 2.0, 1.4])
def test_compiled_ground_view_steps_match_geometry(scale):
    """The step shifts compiled into the library (walked instead of the offsets table at
    scale 1) equal the first steps of the ground-view geometry at every scale."""
    if not os.path.exists(rt.LIBRARY):
        pytest.skip('library not built')
    import numpy as np
    from solweig_amd.sky import gvf_geometry
    off, _, _, second = gvf_geometry(scale)
    n = min(second, 22)
    out = np.zeros((18, n, 2), np.int32)
    rt.call('sw_gvf_fixed_steps', out.ctypes.data, n)
    np.testing.assert_array_equal(out, off[:, :n])
    with pytest.raises(RuntimeError):
        rt.call('sw_gvf_fixed_steps', out.ctypes.data, 23)


def test_bindings_refuse_a_library_of_another_record_size(monkeypatch):
    """A library built for another sw_domain layout (a stale .so beside newer bindings) is refused
    at load time instead of reading domain records at the wrong stride."""
    if not os.path.exists(rt.LIBRARY):
        pytest.skip('library not built')

    class Wider(ctypes.Structure):
        _fields_ = rt.SwDomain._fields_ + [('extra', ctypes.c_int64)]

    monkeypatch.setattr(rt, '_lib', None)
    monkeypatch.setattr(rt, 'SwDomain', Wider)
    with pytest.raises(rt.SolweigHipError, match='336-byte .* records, these bindings use 344'):
        rt.lib()


@pytest.mark.parametrize('rows,cols,real_bytes', [(2048, 1025, 4), (1025, 1024, 8), (4000, 4000, 4)])
def test_noalloc_gvf_refuses_rasters_beyond_its_scratch(rows, cols, real_bytes):
    """The size check runs on the host before any device call (no GPU needed): 64 MiB of samples
    hold 2048 x 1024 pixels in FP32 and 1024 x 1024 in FP64."""
    if not os.path.exists(rt.LIBRARY):
        pytest.skip('library not built')
    L = rt.lib()
    f = L.sw_gvf_noalloc
    vp, ci = ctypes.c_void_p, ctypes.c_int
    f.argtypes = [vp, vp, ci, ci, ci, vp, vp, ci, ci, ci, vp, ci, ci, vp, vp]
    f.restype = ci
    dummy = ctypes.create_string_buffer(64)
    ptrs = (ctypes.c_void_p * 10)(*([ctypes.addressof(dummy)] * 10))
    counts = (ctypes.c_longlong * 10)(*([1] * 10))
    params = (ctypes.c_double * 6)()
    rc = f(ptrs, counts, 10, rows, cols, ctypes.addressof(dummy), ctypes.addressof(dummy), 22, 1, 22, params, 6,
           real_bytes, ctypes.addressof(dummy), None)
    assert rc == -3                                            # SW_ERR_LIMIT
    assert b'scratch holds 67108864' in L.sw_last_error()


def test_compiled_spa_terms_match_the_tables():
    """The SPA term tables compiled into the library (csrc/spa_terms.inc) are the ones of
    spa_tables.py, the tables the host restatement uses."""
    import importlib.util
    csrc = os.path.join(os.path.dirname(rt.__file__), 'csrc')
    spec = importlib.util.spec_from_file_location('gen_spa_terms', os.path.join(csrc, 'gen_spa_terms.py'))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(csrc, 'spa_terms.inc')) as fh:
        assert fh.read() == gen.render()


@pytest.mark.parametrize('groups', [0, 1, 9])
def test_capped_pixel_pass_refuses_caps_out_of_range(groups):
    """sw_day_pixels_capped checks its cap on the host before any device call (no GPU needed)."""
    if not os.path.exists(rt.LIBRARY):
        pytest.skip('library not built')
    L = rt.lib()
    dummy = ctypes.create_string_buffer(64)
    a = ctypes.addressof(dummy)
    rc = L.sw_day_pixels_capped(a, a, 1, 1, 0, 24, a, a, 1, 22, groups, None)
    assert rc == -1                                            # SW_ERR_ARGUMENT
    assert b'sw_day_pixels_capped: bad arguments' in L.sw_last_error()
