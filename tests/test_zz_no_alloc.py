"""sw_gvf_noalloc (the ground-view call without a work buffer) against sw_gvf (GPU): identical
planes on two calls in a row, the second with a smaller raster, in FP32 and FP64, with and
without the land-cover water rule."""
import numpy as np
import pytest

from test_gpu_submodels import _gvf_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    from solweig_amd import runtime
    if runtime.device_count() < 1:
        pytest.skip('no GPU')
    return runtime


@pytest.mark.parametrize('landcover', [0, 1])
@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_noalloc_gvf_equals_the_work_buffer_call(rt, dtype, landcover):
    from solweig_amd import model
    for k, (rows, cols) in enumerate(((181, 150), (61, 47))):          # the second call is smaller
        print(f'sw_gvf_noalloc {np.dtype(dtype).name} landcover {landcover}: {rows}x{cols}', flush=True)
        f = _gvf_fields(70 + k + 2 * landcover, rows, cols, landcover)
        geo = model.prepare_gvf_geometry(f['walls'], f['buildings'], f['dirwalls'], 1.0, 1.0, 22.0)
        args = (geo, f['wallsun'], f['shadow'], f['Tg'], f['Tgwall'], f['Ta'], f['emis'], 0.9, f['alb'],
                5.67e-8, 0.2, f['Twater'], f['lc'], landcover)
        a = model.gvf(*args, dtype=dtype)
        b = model.gvf(*args, dtype=dtype, no_alloc=True)
        for i, (x, y) in enumerate(zip(a.outputs + (a.surface_temperature,), b.outputs + (b.surface_temperature,))):
            xa, ya = x.to_host(), y.to_host()
            assert xa.dtype == np.dtype(dtype) and xa.shape == (rows, cols)
            np.testing.assert_array_equal(xa, ya, err_msg=f'plane {i}')
        assert np.isfinite(a.outputs[0].to_host()).all()
