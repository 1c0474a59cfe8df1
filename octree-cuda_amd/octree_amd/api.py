"""ctypes binding of libOctreeApi.so, the C ABI in include/octreeApi.h.

``Session`` mirrors the reference's call sequence (src/main.cpp:55-109): create the session, set
the cloud description and buffer, configure chunking and subsampling, init, generate, export.
Every call raises ``OctreeError`` with the library's message when the C function returns -1.
There is no fallback: without the built library (or without a GPU) the calls fail.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from ._lib import load_library

OUT_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("r", "<u2"), ("g", "<u2"), ("b", "<u2")])
HIER_DTYPE = np.dtype([("type", "u1"), ("mask", "u1"), ("points", "<u4"), ("offset", "<u8"), ("size", "<u8")])

# (name, restype, argtypes) of every exported entry point
_vp = ctypes.c_void_p
_u8, _u32, _u64, _f, _d, _b = (ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double,
                               ctypes.c_bool)
SIGNATURES = {
    "ocpi_create_session": [ctypes.POINTER(_vp), ctypes.c_int],
    "ocpi_destroy_session": [_vp],
    "ocpi_set_logging_level": [ctypes.c_int],
    "ocpi_set_cloud_type": [_vp, _u8],
    "ocpi_set_cloud_point_amount": [_vp, _u32],
    "ocpi_set_cloud_data_stride": [_vp, _u32],
    "ocpi_set_cloud_scale": [_vp, _d, _d, _d],
    "ocpi_set_cloud_offset": [_vp, _d, _d, _d],
    "ocpi_set_cloud_bb": [_vp, _d, _d, _d, _d, _d, _d],
    "ocpi_set_point_cloud_host": [_vp, _vp],
    "ocpi_configure_chunking": [_vp, _u32, _u32, _f],
    "ocpi_configure_subsampling": [_vp, _u32, _b, _b, _b, _b, _b],
    "ocpi_export_memory_report": [_vp, ctypes.c_char_p],
    "ocpi_export_distribution_histogram": [_vp, ctypes.c_char_p, _u32],
    "ocpi_export_json_report": [_vp, ctypes.c_char_p],
    "ocpi_export_potree": [_vp, ctypes.c_char_p],
    "ocpi_init_octree": [_vp],
    "ocpi_generate_octree": [_vp],
    "ocpi_set_point_cloud_device": [_vp, _vp],
    "ocpi_set_random_seed": [_vp, _u64],
    "ocpi_get_last_error": [ctypes.c_char_p, _u32],
    "ocpi_get_build_summary": [_vp, _vp, _u32],
    "ocpi_copy_nodes": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ocpi_copy_point_lut": [_vp, _vp, _u64],
    "ocpi_copy_output": [_vp, _vp, _u64],
    "ocpi_get_stage_timings": [_vp, _vp, _u32],
    "ocpi_create_service": [ctypes.POINTER(_vp), ctypes.c_int, _u32, _u64],
    "ocpi_create_batch_service": [ctypes.POINTER(_vp), ctypes.c_int, _u32, _u32, _u64],
    "ocpi_destroy_service": [_vp],
    "ocpi_service_submit": [_vp, _vp],
    "ocpi_service_wait": [_vp, _vp, _u32],
    "ocpi_service_set_active_slots": [_vp, _u32],
    "ocpi_service_get_stats": [_vp, _vp, _u32],
    "ocpi_service_reset_stats": [_vp],
    "ocpi_get_request_timing": [_vp, _vp, _u32],
    "ocpi_estimate_footprint": [_vp, _vp],
    "ocpi_compute_bounding_box": [_vp, _u32, _u32, _u8, ctypes.c_int, _vp, _vp],
    "ocpi_compute_bounding_box_noalloc": [_vp, _u32, _u32, _u8, ctypes.c_int, _vp, _vp],
    "ocpi_create_batch_builder": [ctypes.POINTER(_vp), ctypes.c_int],
    "ocpi_destroy_batch_builder": [_vp],
    "ocpi_batch_build": [_vp, _vp, _u32, _vp],
    "ocpi_batch_set_launch": [_vp, _u32, _u32, _u32],
    "ocpi_batch_use_graphs": [_vp, _u32],
    "ocpi_batch_get_timings": [_vp, _vp, _u32],
    "ocpi_estimate_batch_footprint": [_vp, _u32, _vp],
}

BATCH_ENTRY_POINTS = ["ocpi_create_batch_builder", "ocpi_destroy_batch_builder", "ocpi_batch_build",
                      "ocpi_batch_set_launch", "ocpi_batch_use_graphs", "ocpi_batch_get_timings",
                      "ocpi_estimate_batch_footprint"]

SERVICE_ENTRY_POINTS = ["ocpi_create_service", "ocpi_create_batch_service", "ocpi_destroy_service", "ocpi_service_submit", "ocpi_service_wait",
                        "ocpi_service_set_active_slots", "ocpi_service_get_stats", "ocpi_service_reset_stats",
                        "ocpi_get_request_timing", "ocpi_estimate_footprint"]

REFERENCE_ENTRY_POINTS = [n for n in SIGNATURES if n not in [
    "ocpi_set_point_cloud_device", "ocpi_set_random_seed", "ocpi_get_last_error", "ocpi_get_build_summary",
    "ocpi_copy_nodes", "ocpi_copy_point_lut", "ocpi_copy_output", "ocpi_get_stage_timings",
    "ocpi_compute_bounding_box", "ocpi_compute_bounding_box_noalloc"] + SERVICE_ENTRY_POINTS + BATCH_ENTRY_POINTS]

SUMMARY_FIELDS = ("sparse_nodes", "points", "capacity", "used", "internal_nodes", "leaf_nodes", "depth",
                  "dense_nodes", "exported_nodes")
TIMING_FIELDS = ("count_ms", "merge_ms", "distribute_ms", "subsample_ms", "build_ms", "insert_ms", "emit_ms")
SERVICE_STAT_FIELDS = ("submitted", "completed", "failed", "queued", "running", "max_queue_depth", "p50_ms", "p99_ms",
                       "mean_ms", "reserved_bytes", "budget_bytes", "active_slots", "slots", "mean_wait_ms",
                       "max_ms", "peak_reserved_bytes", "setup_ms", "generate_ms", "capture_ms", "batches",
                       "requests_per_batch")

_api = None


def library() -> ctypes.CDLL:
    global _api
    if _api is None:
        lib = load_library("libOctreeApi.so")
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _api = lib
    return _api


class OctreeError(RuntimeError):
    pass


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    library().ocpi_get_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def _check(rc: int, what: str):
    if rc != 0:
        raise OctreeError(f"{what} failed: {last_error()}")


@dataclasses.dataclass
class Nodes:
    key: np.ndarray
    count: np.ndarray
    point_count: np.ndarray
    data_idx: np.ndarray
    parent: np.ndarray
    children: np.ndarray
    flags: np.ndarray

    @property
    def finished(self):
        return (self.flags & 1).astype(bool)

    @property
    def internal(self):
        return (self.flags & 2).astype(bool)


class Session:
    """One octree build session on a HIP device (the reference's void* handle)."""

    def __init__(self, device: int = 0):
        self._lib = library()
        h = _vp()
        _check(self._lib.ocpi_create_session(ctypes.byref(h), device), "ocpi_create_session")
        self.handle = h
        self._keep = None

    def close(self):
        if self.handle:
            _check(self._lib.ocpi_destroy_session(self.handle), "ocpi_destroy_session")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration (src/main.cpp:84-99) --
    def set_cloud(self, cloud, points: int, stride: int, cloud_type: int, cubic_bb, scale, offset=None,
                  device_pointer: int | None = None):
        L, h = self._lib, self.handle
        _check(L.ocpi_set_cloud_type(h, cloud_type), "ocpi_set_cloud_type")
        _check(L.ocpi_set_cloud_point_amount(h, points), "ocpi_set_cloud_point_amount")
        _check(L.ocpi_set_cloud_data_stride(h, stride), "ocpi_set_cloud_data_stride")
        _check(L.ocpi_set_cloud_scale(h, *scale), "ocpi_set_cloud_scale")
        off = offset if offset is not None else cubic_bb[:3]
        _check(L.ocpi_set_cloud_offset(h, *off), "ocpi_set_cloud_offset")
        _check(L.ocpi_set_cloud_bb(h, *cubic_bb), "ocpi_set_cloud_bb")
        if device_pointer is not None:
            _check(L.ocpi_set_point_cloud_device(h, device_pointer), "ocpi_set_point_cloud_device")
            self._keep = cloud
        else:
            arr = np.ascontiguousarray(cloud, dtype=np.uint8)
            self._keep = arr
            _check(L.ocpi_set_point_cloud_host(h, arr.ctypes.data), "ocpi_set_point_cloud_host")

    def configure(self, spec):
        """spec: workloads.BuildSpec (chunking + subsampling configuration and seed)."""
        L, h = self._lib, self.handle
        _check(L.ocpi_configure_chunking(h, spec.chunking_grid, spec.merging_threshold, spec.output_factor),
               "ocpi_configure_chunking")
        _check(L.ocpi_configure_subsampling(h, spec.subsampling_grid, spec.intra, spec.inter, spec.replacement,
                                            spec.random, spec.weighted), "ocpi_configure_subsampling")
        _check(L.ocpi_set_random_seed(h, spec.seed), "ocpi_set_random_seed")

    def init(self):
        _check(self._lib.ocpi_init_octree(self.handle), "ocpi_init_octree")

    def generate(self):
        _check(self._lib.ocpi_generate_octree(self.handle), "ocpi_generate_octree")

    # -- exports --
    def export_potree(self, directory: str):
        _check(self._lib.ocpi_export_potree(self.handle, directory.encode()), "ocpi_export_potree")

    def export_json_report(self, filename: str):
        _check(self._lib.ocpi_export_json_report(self.handle, filename.encode()), "ocpi_export_json_report")

    def export_distribution_histogram(self, filename: str, bin_width: int = 0):
        _check(self._lib.ocpi_export_distribution_histogram(self.handle, filename.encode(), bin_width),
               "ocpi_export_distribution_histogram")

    def export_memory_report(self, filename: str):
        _check(self._lib.ocpi_export_memory_report(self.handle, filename.encode()), "ocpi_export_memory_report")

    # -- results --
    def summary(self) -> dict:
        v = np.zeros(len(SUMMARY_FIELDS), np.uint64)
        _check(self._lib.ocpi_get_build_summary(self.handle, v.ctypes.data, len(v)), "ocpi_get_build_summary")
        return {k: int(x) for k, x in zip(SUMMARY_FIELDS, v)}

    def timings(self) -> dict:
        v = np.zeros(len(TIMING_FIELDS), np.float32)
        _check(self._lib.ocpi_get_stage_timings(self.handle, v.ctypes.data, len(v)), "ocpi_get_stage_timings")
        return {k: float(x) for k, x in zip(TIMING_FIELDS, v)}

    def nodes(self) -> Nodes:
        m = self.summary()["sparse_nodes"]
        n = Nodes(np.empty(m, np.uint32), np.empty(m, np.uint32), np.empty(m, np.uint32), np.empty(m, np.uint64),
                  np.empty(m, np.int32), np.empty((m, 8), np.int32), np.empty(m, np.uint8))
        _check(self._lib.ocpi_copy_nodes(self.handle, n.key.ctypes.data, n.count.ctypes.data,
                                         n.point_count.ctypes.data, n.data_idx.ctypes.data, n.parent.ctypes.data,
                                         n.children.ctypes.data, n.flags.ctypes.data), "ocpi_copy_nodes")
        return n

    def point_lut(self, count: int) -> np.ndarray:
        out = np.empty(count, np.uint32)
        _check(self._lib.ocpi_copy_point_lut(self.handle, out.ctypes.data, count), "ocpi_copy_point_lut")
        return out

    def output(self, count: int) -> np.ndarray:
        out = np.empty(count, OUT_DTYPE)
        _check(self._lib.ocpi_copy_output(self.handle, out.ctypes.data, count), "ocpi_copy_output")
        return out

    def footprint(self) -> int:
        v = ctypes.c_uint64()
        _check(self._lib.ocpi_estimate_footprint(self.handle, ctypes.byref(v)), "ocpi_estimate_footprint")
        return int(v.value)

    def request_timing(self) -> dict:
        v = np.zeros(3, np.float64)
        _check(self._lib.ocpi_get_request_timing(self.handle, v.ctypes.data, 3), "ocpi_get_request_timing")
        return {"queue_ms": float(v[0]), "run_ms": float(v[1]), "total_ms": float(v[2])}


class Service:
    """Continuous batching of builds: configured sessions are requests. The default is the slot
    service (ocpi_create_service: `slots` single-cloud builders); with batch_size the batch
    service (ocpi_create_batch_service: `slots` workers that each build up to batch_size queued
    requests together)."""

    def __init__(self, device: int = 0, slots: int = 4, memory_budget: int = 200 << 30,
                 batch_size: int | None = None):
        self._lib = library()
        h = _vp()
        if batch_size is None:
            _check(self._lib.ocpi_create_service(ctypes.byref(h), device, slots, memory_budget),
                   "ocpi_create_service")
        else:
            _check(self._lib.ocpi_create_batch_service(ctypes.byref(h), device, slots, batch_size, memory_budget),
                   "ocpi_create_batch_service")
        self.handle = h

    def close(self):
        if self.handle:
            _check(self._lib.ocpi_destroy_service(self.handle), "ocpi_destroy_service")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, session: Session):
        _check(self._lib.ocpi_service_submit(self.handle, session.handle), "ocpi_service_submit")

    def wait(self, session: Session, timeout_ms: int = 600_000):
        _check(self._lib.ocpi_service_wait(self.handle, session.handle, timeout_ms), "ocpi_service_wait")

    def set_active_slots(self, slots: int):
        _check(self._lib.ocpi_service_set_active_slots(self.handle, slots), "ocpi_service_set_active_slots")

    def stats(self) -> dict:
        v = np.zeros(len(SERVICE_STAT_FIELDS), np.float64)
        _check(self._lib.ocpi_service_get_stats(self.handle, v.ctypes.data, len(v)), "ocpi_service_get_stats")
        return {k: float(x) for k, x in zip(SERVICE_STAT_FIELDS, v)}

    def reset_stats(self):
        _check(self._lib.ocpi_service_reset_stats(self.handle), "ocpi_service_reset_stats")


BATCH_TIMING_FIELDS = ("count_ms", "merge_ms", "distribute_ms", "subsample_ms", "device_ms", "host_ms", "copy_ms")


class BatchBuilder:
    """Batched builds (ocpi_create_batch_builder): configured sessions built together, one launch
    per stage for all clouds of a group; each session then holds its result."""

    def __init__(self, device: int = 0):
        self._lib = library()
        h = _vp()
        _check(self._lib.ocpi_create_batch_builder(ctypes.byref(h), device), "ocpi_create_batch_builder")
        self.handle = h

    def close(self):
        if self.handle:
            _check(self._lib.ocpi_destroy_batch_builder(self.handle), "ocpi_destroy_batch_builder")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, sessions, raise_on_failure: bool = True) -> np.ndarray:
        """Builds the sessions; returns the per-session status (0 ok, -1 failed)."""
        handles = (_vp * max(1, len(sessions)))(*[s.handle for s in sessions])
        status = np.zeros(max(1, len(sessions)), np.int32)
        rc = self._lib.ocpi_batch_build(self.handle, handles, len(sessions), status.ctypes.data)
        if rc != 0 and raise_on_failure:
            raise OctreeError(f"ocpi_batch_build failed: {last_error()}")
        return status[:len(sessions)]

    def set_launch(self, thread_grid: int, tile_grid: int, sort_grid: int):
        _check(self._lib.ocpi_batch_set_launch(self.handle, thread_grid, tile_grid, sort_grid),
               "ocpi_batch_set_launch")

    def use_graphs(self, on: bool):
        _check(self._lib.ocpi_batch_use_graphs(self.handle, 1 if on else 0), "ocpi_batch_use_graphs")

    def timings(self) -> dict:
        v = np.zeros(len(BATCH_TIMING_FIELDS), np.float32)
        _check(self._lib.ocpi_batch_get_timings(self.handle, v.ctypes.data, len(v)), "ocpi_batch_get_timings")
        return {k: float(x) for k, x in zip(BATCH_TIMING_FIELDS, v)}


def batch_footprint(sessions) -> int:
    handles = (_vp * max(1, len(sessions)))(*[s.handle for s in sessions])
    v = ctypes.c_uint64()
    _check(library().ocpi_estimate_batch_footprint(handles, len(sessions), ctypes.byref(v)),
           "ocpi_estimate_batch_footprint")
    return int(v.value)


def set_logging_level(level: int):
    _check(library().ocpi_set_logging_level(level), "ocpi_set_logging_level")
