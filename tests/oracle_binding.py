"""ctypes binding of the CPU oracle (oracle/octree_oracle.cpp). Test infrastructure only."""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "liboctree_oracle.so")

OUT_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("r", "<u2"), ("g", "<u2"), ("b", "<u2")])
HIER_DTYPE = np.dtype([("type", "u1"), ("mask", "u1"), ("points", "<u4"), ("offset", "<u8"), ("size", "<u8")])
assert OUT_DTYPE.itemsize == 18 and HIER_DTYPE.itemsize == 22


class OracleConfig(ctypes.Structure):
    _fields_ = [
        ("point_amount", ctypes.c_uint32),
        ("data_stride", ctypes.c_uint32),
        ("cloud_type", ctypes.c_uint32),
        ("bb_min", ctypes.c_double * 3),
        ("bb_max", ctypes.c_double * 3),
        ("scale", ctypes.c_double * 3),
        ("chunking_grid", ctypes.c_uint32),
        ("merging_threshold", ctypes.c_uint32),
        ("output_factor", ctypes.c_float),
        ("subsampling_grid", ctypes.c_uint32),
        ("intra", ctypes.c_uint32),
        ("inter", ctypes.c_uint32),
        ("replacement", ctypes.c_uint32),
        ("random", ctypes.c_uint32),
        ("weighted", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("threads", ctypes.c_uint32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        vp = ctypes.c_void_p
        L.oracle_build.argtypes = [vp, ctypes.POINTER(OracleConfig), ctypes.POINTER(ctypes.c_int)]
        L.oracle_build.restype = vp
        L.oracle_free.argtypes = [vp]
        L.oracle_node_amount.argtypes = [vp]
        L.oracle_node_amount.restype = ctypes.c_uint32
        L.oracle_get_nodes.argtypes = [vp] * 8
        L.oracle_capacity.argtypes = [vp]
        L.oracle_capacity.restype = ctypes.c_uint64
        L.oracle_used_points.argtypes = [vp]
        L.oracle_used_points.restype = ctypes.c_uint64
        L.oracle_get_lut.argtypes = [vp, vp]
        L.oracle_get_output.argtypes = [vp, vp]
        L.oracle_hierarchy_size.argtypes = [vp]
        L.oracle_hierarchy_size.restype = ctypes.c_uint64
        L.oracle_get_hierarchy.argtypes = [vp, vp]
        L.oracle_get_summary.argtypes = [vp, vp]
        L.oracle_subsample_node.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, vp]
        L.oracle_subsample_node.restype = ctypes.c_uint32
        L.oracle_map_point_to_grid.argtypes = [ctypes.c_double] * 3 + [ctypes.c_uint32] + [ctypes.c_double] * 4
        L.oracle_map_point_to_grid.restype = ctypes.c_uint32
        L.oracle_encode_colors.argtypes = [ctypes.c_uint16] * 3
        L.oracle_encode_colors.restype = ctypes.c_uint64
        L.oracle_decode_colors.argtypes = [ctypes.c_uint64, vp]
        L.oracle_output_capacity.argtypes = [ctypes.c_uint32, ctypes.c_float]
        L.oracle_output_capacity.restype = ctypes.c_uint64
        L.oracle_random_rank.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_random_rank.restype = ctypes.c_uint32
        L.oracle_level_tables.argtypes = [ctypes.c_uint32, vp, vp, vp, vp]
        L.oracle_level_tables.restype = ctypes.c_uint32
        L.oracle_node_statistics.argtypes = [ctypes.c_uint32, vp, vp, vp, ctypes.c_uint32, vp]
        _lib = L
    return _lib


@dataclasses.dataclass
class OracleResult:
    status: int
    key: np.ndarray = None        # dense key per sparse node
    count: np.ndarray = None      # dense subtree count
    point_count: np.ndarray = None
    data_idx: np.ndarray = None
    parent: np.ndarray = None
    children: np.ndarray = None   # (M, 8)
    flags: np.ndarray = None      # bit0 finished, bit1 internal
    lut: np.ndarray = None
    out: np.ndarray = None
    hierarchy: bytes = b""
    capacity: int = 0
    used: int = 0
    summary: np.ndarray = None    # exportedNodes, exportedPoints, internalNodes, referenceBoundBinds, maxLeaf
    handle: object = None         # Handle when built with keep=True

    @property
    def finished(self):
        return (self.flags & 1).astype(bool)

    @property
    def internal(self):
        return (self.flags & 2).astype(bool)


def make_config(n, stride, cloud_type, cubic, scale, build, threads=0) -> OracleConfig:
    c = OracleConfig()
    c.point_amount = n
    c.data_stride = stride
    c.cloud_type = cloud_type
    for a in range(3):
        c.bb_min[a] = cubic[a]
        c.bb_max[a] = cubic[3 + a]
        c.scale[a] = scale[a]
    c.chunking_grid = build.chunking_grid
    c.merging_threshold = build.merging_threshold
    c.output_factor = build.output_factor
    c.subsampling_grid = build.subsampling_grid
    c.intra = int(build.intra)
    c.inter = int(build.inter)
    c.replacement = int(build.replacement)
    c.random = int(build.random)
    c.weighted = int(build.weighted)
    c.seed = build.seed
    c.threads = threads
    return c


class Handle:
    """A live oracle build (for per-node recomputation); keeps the cloud buffer alive."""

    def __init__(self, handle, cloud):
        self.handle = handle
        self.cloud = cloud

    def subsample_node(self, node_id, idx, rgb):
        n = len(idx)
        idx = np.ascontiguousarray(idx, np.uint32)
        rgb = np.ascontiguousarray(rgb, np.uint16).reshape(-1)
        out_idx = np.empty(max(n, 1), np.uint32)
        out_rgb = np.empty(3 * max(n, 1), np.uint16)
        k = lib().oracle_subsample_node(self.handle, self.cloud.ctypes.data, node_id, n, idx.ctypes.data,
                                        rgb.ctypes.data, out_idx.ctypes.data, out_rgb.ctypes.data)
        return out_idx[:k], out_rgb[:3 * k].reshape(k, 3)

    def close(self):
        if self.handle:
            lib().oracle_free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def build(cloud: np.ndarray, n, stride, cloud_type, cubic, scale, spec, threads=0, want_data=True,
          keep=False) -> OracleResult:
    L = lib()
    cfg = make_config(n, stride, cloud_type, cubic, scale, spec, threads)
    status = ctypes.c_int(0)
    h = L.oracle_build(cloud.ctypes.data, ctypes.byref(cfg), ctypes.byref(status))
    if not h:
        return OracleResult(status.value)
    try:
        m = L.oracle_node_amount(h)
        r = OracleResult(status.value)
        r.key = np.empty(m, np.uint32)
        r.count = np.empty(m, np.uint32)
        r.point_count = np.empty(m, np.uint32)
        r.data_idx = np.empty(m, np.uint64)
        r.parent = np.empty(m, np.int32)
        r.children = np.empty((m, 8), np.int32)
        r.flags = np.empty(m, np.uint8)
        L.oracle_get_nodes(h, r.key.ctypes.data, r.count.ctypes.data, r.point_count.ctypes.data,
                           r.data_idx.ctypes.data, r.parent.ctypes.data, r.children.ctypes.data,
                           r.flags.ctypes.data)
        r.capacity = L.oracle_capacity(h)
        r.used = L.oracle_used_points(h)
        r.summary = np.empty(5, np.uint64)
        L.oracle_get_summary(h, r.summary.ctypes.data)
        if status.value == 0 and want_data:
            r.lut = np.empty(r.capacity, np.uint32)
            L.oracle_get_lut(h, r.lut.ctypes.data)
            r.out = np.empty(r.capacity, OUT_DTYPE)
            L.oracle_get_output(h, r.out.ctypes.data)
            hs = L.oracle_hierarchy_size(h)
            hb = np.empty(hs, np.uint8)
            L.oracle_get_hierarchy(h, hb.ctypes.data)
            r.hierarchy = hb.tobytes()
        if keep:
            r.handle = Handle(h, cloud)
            h = None
        return r
    finally:
        if h:
            L.oracle_free(h)


def level_tables(grid):
    L = lib()
    depth = ctypes.c_uint32()
    grids = np.zeros(10, np.uint32)
    nodes = np.zeros(10, np.uint32)
    offs = np.zeros(10, np.uint32)
    dense = L.oracle_level_tables(grid, ctypes.byref(depth), grids.ctypes.data, nodes.ctypes.data, offs.ctypes.data)
    d = depth.value
    return d, grids[:d + 1], nodes[:d + 1], offs[:d + 1], dense


def map_point_to_grid(x, y, z, grid, bb_size, mn):
    return lib().oracle_map_point_to_grid(x, y, z, grid, bb_size, mn[0], mn[1], mn[2])


def encode_colors(r, g, b):
    return lib().oracle_encode_colors(r, g, b)


def decode_colors(e):
    rgb = np.zeros(3, np.uint16)
    lib().oracle_decode_colors(e, rgb.ctypes.data)
    return tuple(int(v) for v in rgb)


def output_capacity(n, factor):
    return lib().oracle_output_capacity(n, factor)


def random_rank(seed, node_key, point_idx):
    return lib().oracle_random_rank(seed, node_key, point_idx)


def node_statistics(point_count, children, flags, root):
    st = np.zeros(8, np.uint64)
    lib().oracle_node_statistics(len(point_count), np.ascontiguousarray(point_count, np.uint32).ctypes.data,
                                 np.ascontiguousarray(children, np.int32).ctypes.data,
                                 np.ascontiguousarray(flags, np.uint8).ctypes.data, root, st.ctypes.data)
    mean = np.array([st[5]], np.uint32).view(np.float32)[0]
    stdev = np.array([st[6]], np.uint32).view(np.float32)[0]
    return dict(leaves=int(st[0]), internal=int(st[1]), min=int(st[2]), max=int(st[3]), max_leaf_depth=int(st[4]),
                mean=float(mean), stdev=float(stdev), point_sum=int(st[7]))
