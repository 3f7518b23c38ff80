"""The no-allocation dot product (fem_dot_device_noalloc) against fem_dot,
through the C ABI: bit-identical results (same kernels, same summation order)
on consecutive calls of decreasing size, which share the device's static
scratch array, and the size limit's error."""
import ctypes as C

import numpy as np
import pytest

import femgpu as F

pytestmark = pytest.mark.gpu


def test_dot_noalloc_matches_fem_dot():
    L = F.lib()
    rng = np.random.default_rng(5)
    dout = F.DeviceArray(1)
    for n in (3_000_017, 12_345, 1):
        print(f"dot_noalloc n={n}", flush=True)
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        da, db = F.DeviceArray.from_host(a), F.DeviceArray.from_host(b)
        rc = L.fem_dot_device_noalloc(n, da.ptr, db.ptr, dout.ptr, None)
        assert rc == 0, L.fem_last_error()
        F.device_synchronize()
        got = dout.download()[0]
        ref = F.dot(a, b)
        assert got == ref, (n, got, ref)
        assert abs(got - float(np.dot(a, b))) <= 1e-12 * np.abs(a * b).sum()
    # beyond the limit: rejected before anything is launched
    rc = L.fem_dot_device_noalloc(1 << 32, da.ptr, db.ptr, dout.ptr, None)
    assert rc == 1 and b"2^32" in L.fem_last_error()
    rc = L.fem_dot_device_noalloc(10, None, db.ptr, dout.ptr, None)
    assert rc == 1
