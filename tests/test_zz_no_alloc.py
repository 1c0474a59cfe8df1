"""ocpi_compute_bounding_box_noalloc (no device allocation: partials in a static device array)
against ocpi_compute_bounding_box on the same device-resident clouds, bit for bit, on two calls in
a row (the second with a smaller cloud, so the array still holds the first call's partials) and
for both record types."""
import ctypes

import pytest

from hip_memory import DeviceBuffer
from octree_amd import api
from parity import synthetic

pytestmark = pytest.mark.gpu


def _box(fn, buf, n, stride, cloud_type):
    real = (ctypes.c_double * 6)()
    cubic = (ctypes.c_double * 6)()
    rc = fn(buf.ptr, n, stride, cloud_type, 0, real, cubic)
    assert rc == 0, api.last_error()
    return list(real), list(cubic)


@pytest.mark.parametrize("cloud_type", [0, 1])
def test_noalloc_bounding_box_equals_allocating_one(cloud_type):
    lib = api.library()
    for n in (900_001, 70_003):  # the second call is smaller
        print(f"case bounding box noalloc type={cloud_type} n={n}", flush=True)
        cloud, spec, _ = synthetic(n, cloud_type, box=(40.0, 9.0, 3.0), seed=50 + n % 7,
                                   offset=0.0 if cloud_type == 0 else 7e5)
        buf = DeviceBuffer(cloud)
        try:
            want = _box(lib.ocpi_compute_bounding_box, buf, n, spec.stride, cloud_type)
            got = _box(lib.ocpi_compute_bounding_box_noalloc, buf, n, spec.stride, cloud_type)
        finally:
            buf.close()
        assert got == want


def test_noalloc_bounding_box_rejects_an_empty_cloud():
    real = (ctypes.c_double * 6)()
    cubic = (ctypes.c_double * 6)()
    assert api.library().ocpi_compute_bounding_box_noalloc(None, 0, 15, 0, 0, real, cubic) == -1
    assert "empty" in api.last_error()
