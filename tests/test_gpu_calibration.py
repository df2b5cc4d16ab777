"""The HBM read calibration behind bench.py's `stream_read_gbs`: called through the C ABI
on device memory allocated with the HIP runtime, it returns a plausible read rate."""
import ctypes as C

import pytest

from ceos_amd import capi

pytestmark = pytest.mark.gpu


def test_read_bandwidth_on_device():
    hip = C.CDLL("libamdhip64.so")
    x, out = C.c_void_p(), C.c_void_p()
    n = 1 << 24  # 64 MiB of floats
    n_out = 256 * 256
    assert hip.hipMalloc(C.byref(x), C.c_size_t(n * 4)) == 0
    try:
        assert hip.hipMalloc(C.byref(out), C.c_size_t(n_out * 4)) == 0
        assert hip.hipMemset(x, 0, C.c_size_t(n * 4)) == 0
        g = C.c_double()
        capi.check(capi.lib().ceos_read_bandwidth(x, n, out, n_out, 3, None, C.byref(g)))
        assert 50.0 < g.value < 20000.0, g.value
    finally:
        hip.hipFree(x)
        if out.value:
            hip.hipFree(out)
