"""swm_gpu_sums_noalloc (the allocation-free variant of swm_gpu_sums, whose
partials, result and cos(lat) rows sit in the library's static device array)
called through the C-ABI: bit-identical to swm_gpu_sums on two calls in a row,
the second on a smaller grid, in FP64 and FP32, with and without a tracer;
grids beyond its row limit are refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import same
from swm_amd import configs, gpu

pytestmark = pytest.mark.gpu


def via_abi(L, m, fn, tracer):
    out = (C.c_double * 3)()
    rc = getattr(L, fn)(m._h, int(tracer), C.cast(out, C.POINTER(C.c_double)))
    assert rc == 0, L.swm_gpu_last_error(m._h).decode()
    return np.array(out[:])


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP32])
def test_noalloc_sums_equal_allocating_sums(gpu_lib, precision):
    L = gpu.load_library()
    big = gpu.SwmGpu(configs.c2_nonlinear(seed=2, N=300, dt=10.0), precision).init_from_setup()
    small = gpu.SwmGpu(configs.c2_nonlinear(seed=3, N=90, dt=10.0, tracer=False), precision).init_from_setup()
    big.step(13)
    small.step(7)
    for m, tr in ((big, 0), (small, -1)):        # the second call on a smaller grid
        a = via_abi(L, m, "swm_gpu_sums_noalloc", tr)
        b = via_abi(L, m, "swm_gpu_sums", tr)
        assert same(a, b), (a, b)
        assert np.isfinite(a[:2]).all() and a[1] > 0.0
    a = via_abi(L, big, "swm_gpu_sums_noalloc", -1)   # the larger grid again, after the smaller one
    assert same(a, via_abi(L, big, "swm_gpu_sums", -1))
    big.close()
    small.close()


def test_noalloc_sums_refuse_too_many_rows(gpu_lib):
    L = gpu.load_library()
    s = configs.basin(8, 16400, 5e3, 5e3, variant=gpu.FULLY_NONLINEAR, H0=100.0, dt=5.0)
    from swm_amd import host
    m = gpu.SwmGpu(host.finalize(s)).init_from_setup()
    out = (C.c_double * 3)()
    rc = L.swm_gpu_sums_noalloc(m._h, -1, C.cast(out, C.POINTER(C.c_double)))
    assert rc == 4 and b"16384" in L.swm_gpu_last_error(m._h)
    m.close()
