"""The C-ABI library without a GPU: it loads, exports exactly the declared entry points, and every
call fails cleanly (-1 and a message, never a process exit) where the reference would crash or
exit. No compute call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from octree_amd import api
from octree_amd import workloads as W
from octree_amd._lib import library_path

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "octreeApi.h")

# OctreeApi/include/octreeApi.h:42-72
REFERENCE_API = {
    "ocpi_create_session", "ocpi_destroy_session", "ocpi_set_logging_level", "ocpi_set_cloud_type",
    "ocpi_set_cloud_point_amount", "ocpi_set_cloud_data_stride", "ocpi_set_cloud_scale", "ocpi_set_cloud_offset",
    "ocpi_set_cloud_bb", "ocpi_set_point_cloud_host", "ocpi_configure_chunking", "ocpi_configure_subsampling",
    "ocpi_export_memory_report", "ocpi_export_distribution_histogram", "ocpi_export_json_report",
    "ocpi_export_potree", "ocpi_init_octree", "ocpi_generate_octree",
}


def declared():
    text = open(HEADER).read()
    return set(re.findall(r"EXPORTED\s+int\s+(\w+)\s*\(", text))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", library_path("libOctreeApi.so")], check=True,
                         capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_the_reference_api():
    assert REFERENCE_API <= declared()
    assert len(REFERENCE_API) == 18


def test_library_exports_exactly_the_declared_symbols():
    assert exported() == declared()


def test_binding_covers_every_symbol():
    assert set(api.SIGNATURES) == declared()
    assert set(api.REFERENCE_ENTRY_POINTS) == REFERENCE_API
    assert set(api.SERVICE_ENTRY_POINTS) <= declared()
    lib = api.library()
    for name in declared():
        assert hasattr(lib, name)


@pytest.mark.parametrize("name", sorted(n for n in REFERENCE_API | {"ocpi_set_random_seed", "ocpi_init_octree"}
                                        if n not in ("ocpi_create_session", "ocpi_set_logging_level")))
def test_null_session_returns_error(name):
    lib = api.library()
    args = []
    for t in api.SIGNATURES[name]:
        if t is ctypes.c_char_p:
            args.append(b"/tmp/x")
        elif t is ctypes.c_void_p:
            args.append(None)
        elif t in (ctypes.c_float, ctypes.c_double):
            args.append(1.0)
        else:
            args.append(1)
    assert getattr(lib, name)(*args) == -1
    assert "Session" in api.last_error()


@pytest.mark.parametrize("name", ["ocpi_destroy_service", "ocpi_service_submit", "ocpi_service_wait",
                                  "ocpi_service_set_active_slots", "ocpi_service_get_stats",
                                  "ocpi_service_reset_stats"])
def test_null_service_returns_error(name):
    lib = api.library()
    args = [None] + [None if t is ctypes.c_void_p else 1 for t in api.SIGNATURES[name][1:]]
    assert getattr(lib, name)(*args) == -1
    assert "BuildService" in api.last_error() or "Session" in api.last_error()


def test_create_service_reports_missing_device():
    lib = api.library()
    h = ctypes.c_void_p()
    assert lib.ocpi_create_service(ctypes.byref(h), 12345, 2, 1 << 30) == -1
    assert "device" in api.last_error()
    assert lib.ocpi_create_service(ctypes.byref(h), 0, 0, 1 << 30) == -1


def test_create_session_reports_missing_device():
    lib = api.library()
    h = ctypes.c_void_p()
    rc = lib.ocpi_create_session(ctypes.byref(h), 12345)
    assert rc == -1 and "device" in api.last_error()
    assert lib.ocpi_create_session(None, 0) == -1


def test_logging_level():
    for level in (0, 1, 2, 3, 9):
        assert api.library().ocpi_set_logging_level(level) == 0
    api.library().ocpi_set_logging_level(2)


def test_generator_is_counter_based():
    spec = W.CloudSpec(10_000, 1, (50.0, 40.0, 10.0), 5, 42, offset=1e6)
    full = W.generate(spec, threads=1)
    part = W.generate(spec, first=3_000, count=4_000, threads=3)
    assert np.array_equal(full[3_000 * 27:7_000 * 27], part)
    assert np.array_equal(full, W.generate(spec, threads=8))


def test_bounding_box_matches_the_cli_formula():
    spec = W.CloudSpec(20_000, 0, (100.0, 60.0, 20.0), 4, 1)
    cloud = W.generate(spec)
    real, cubic = W.bounding_box(cloud, spec.points, 15, 0)
    xyz = cloud.reshape(-1, 15)[:, :12].copy().view(np.float32).reshape(-1, 3).astype(np.float64)
    assert real == list(xyz.min(0)) + list(xyz.max(0))
    dims = xyz.max(0) - xyz.min(0)
    side = dims.max()
    mn = xyz.min(0) - (side - dims) / 2.0
    assert cubic == list(mn) + list(mn + side)
