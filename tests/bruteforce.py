"""Pure-Python brute-force restatement of the canonical octree build, for tiny clouds only.

It is a second, independent formulation used to pin the C++ oracle: instead of dense grids and a
K4-style merge it derives every node from the point set directly (the level-l cell of a point is
its level-0 cell shifted right by l, point_count_propagation.cuh:58-70), finds each leaf as the
first finished ancestor (point_distributing.cuh:56-60), and subsamples with dictionaries keyed by
cell. Floating-point expressions follow the reference's order (kernel_helpers.cuh:11-25,
kernel_inter_cell_filtering.cuh:109-194); Python floats are IEEE doubles, numpy float32 stands in
for the reference's float arithmetic in the weighted filter.
"""
from __future__ import annotations

import math
import struct

import numpy as np

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def fmix32(h):
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    h ^= h >> 16
    return h


def random_rank(seed, node_key, idx):
    salt = mix64((seed + 0x9E3779B97F4A7C15 * (node_key + 1)) & M64) >> 32
    return fmix32(idx ^ salt)


def grid_coords(p, grid, bb_size, mn):
    t = bb_size / grid
    u = [(p[a] - mn[a]) / t for a in range(3)]
    t = grid - 1.0
    return [int(min(v, t)) if not math.isnan(v) else int(t) for v in u]


def encode(r, g, b):
    return ((r << 46) | (g << 28) | (b << 10) | 1) & M64


def decode(e):
    amount = e & 0x3FF
    if amount == 0:
        return (0, 0, 0)
    return tuple(((((e >> s) & 0xFFFF) // amount) & 0xFFFF) for s in (46, 28, 10))


def read_points(cloud: bytes, n, stride, cloud_type):
    fmt = "<3f" if cloud_type == 0 else "<3d"
    cb = 12 if cloud_type == 0 else 24
    pts, cols = [], []
    for i in range(n):
        rec = cloud[i * stride:(i + 1) * stride]
        pts.append(struct.unpack(fmt, rec[:cb]))
        cols.append((rec[cb], rec[cb + 1], rec[cb + 2]))
    return pts, cols


def build(cloud: bytes, n, stride, cloud_type, cubic, scale, cfg):
    """Returns a dict with nodes (sorted by canonical id), lut and out (lists) and used."""
    G = cfg.chunking_grid
    depth = {2 ** d: d for d in range(1, 10)}[G]
    T = cfg.merging_threshold
    pts, cols = read_points(cloud, n, stride, cloud_type)
    side = cubic[3] - cubic[0]
    mn = cubic[:3]

    # level-0 cell of each point, then every ancestor cell by shifting
    cell0 = [grid_coords(p, G, side, mn) for p in pts]
    counts = {}
    for c in cell0:
        for l in range(depth + 1):
            key = (l, c[0] >> l, c[1] >> l, c[2] >> l)
            counts[key] = counts.get(key, 0) + 1

    def linear(key):
        l, x, y, z = key
        g = G >> l
        return x + y * g + z * g * g

    keys = sorted(counts, key=lambda k: (k[0], linear(k)))
    nid = {k: i for i, k in enumerate(keys)}
    M = len(keys)
    root = M - 1
    internal = [k[0] >= 1 and counts[k] >= T for k in keys]

    def parent_key(k):
        return (k[0] + 1, k[1] >> 1, k[2] >> 1, k[3] >> 1)

    def child_keys(k):
        out = []
        for s in range(8):
            ck = (k[0] - 1, 2 * k[1] + ((s >> 2) & 1), 2 * k[2] + ((s >> 1) & 1), 2 * k[3] + (s & 1))
            out.append(ck if (k[0] >= 1 and ck in counts) else None)
        return out

    finished = []
    for i, k in enumerate(keys):
        if i == root:
            finished.append(True)
        else:
            finished.append(internal[nid[parent_key(k)]])
    point_count = [0 if internal[i] else counts[k] for i, k in enumerate(keys)]
    children = [[(nid[c] if c is not None else -1) for c in child_keys(k)] for k in keys]

    # leaf blocks: parents in id order, children in slot order
    data_idx = [0] * M
    cursor = 0
    if not internal[root]:
        data_idx[root] = 0
        cursor = counts[keys[root]]
    else:
        for pid in range(M):
            if not internal[pid]:
                continue
            for c in children[pid]:
                if c != -1 and not internal[c]:
                    data_idx[c] = cursor
                    cursor += point_count[c]
    leaf_offset = cursor

    E = int(np.float32(n) * np.float32(cfg.output_factor))
    lut = [0] * E
    out = [(0, 0, 0, 0, 0, 0)] * E
    fill = {}
    inv = [1.0 / s for s in scale]

    def quant(p):
        return tuple(int(math.floor((p[a] - mn[a]) * inv[a])) for a in range(3))

    for i in range(n):
        k = (0, *cell0[i])
        s = nid[k]
        while not finished[s]:
            s = nid[parent_key(keys[s])]
        pos = data_idx[s] + fill.get(s, 0)
        fill[s] = fill.get(s, 0) + 1
        lut[pos] = i
        out[pos] = (*quant(pts[i]), *cols[i])

    # node minima, top-down (tools.cu:29-61)
    node_min = {root: list(mn)}
    for pid in range(M - 1, -1, -1):
        k = keys[pid]
        if k[0] == 0 or pid not in node_min:
            continue
        side_l = side / (2.0 ** (depth - k[0]))
        cs = side_l / 2.0
        for s, c in enumerate(children[pid]):
            if c == -1:
                continue
            m = list(node_min[pid])
            if s & 4:
                m[0] = m[0] + cs
            if s & 2:
                m[1] = m[1] + cs
            if s & 1:
                m[2] = m[2] + cs
            node_min[c] = m

    g = cfg.subsampling_grid
    mode = "intra" if cfg.intra else ("weighted" if (cfg.inter and cfg.weighted) else ("inter" if cfg.inter else "none"))
    subs = {}

    def subsample(pid):
        k = keys[pid]
        side_l = side / (2.0 ** (depth - k[0]))
        m = node_min[pid]
        bb = (m[0] + side_l) - m[0]
        t = bb / g
        diag = math.pow((t * 3) * (t * 3) * 3, 0.5) / 2.0
        inputs = []
        for c in children[pid]:
            if c == -1:
                continue
            if internal[c]:
                inputs.extend(subs[c])
            else:
                for j in range(point_count[c]):
                    pos = data_idx[c] + j
                    inputs.append((lut[pos], out[pos]))
        occ = {}
        for idx, rec in inputs:
            cc = tuple(grid_coords(pts[idx], g, bb, m))
            rank = random_rank(cfg.seed, _dense_key(G, k), idx) if cfg.random else idx
            occ.setdefault(cc, []).append((rank, idx, rec))
        color = {}
        if mode == "intra":
            for cc, lst in occ.items():
                color[cc] = sum(encode(*rec[3:]) for _, _, rec in lst) & M64
        elif mode in ("inter", "weighted"):
            acc = {cc: (0 if mode == "inter" else [np.float32(0)] * 4) for cc in occ}
            for idx, rec in inputs:
                p = pts[idx]
                ix, iy, iz = grid_coords(p, g, bb, m)
                e = encode(*rec[3:])
                for ox in (-1, 0, 1):
                    for oy in (-1, 0, 1):
                        for oz in (-1, 0, 1):
                            nc = (ix + ox, iy + oy, iz + oz)
                            if nc not in acc:
                                continue
                            if mode == "inter":
                                acc[nc] = (acc[nc] + e) & M64
                            else:
                                uX = (m[0] + nc[0] * t + (t * 0.5)) - p[0]
                                uY = (m[1] + nc[1] * t + (t * 0.5)) - p[1]
                                uZ = (m[2] + nc[2] * t + (t * 0.5)) - p[2]
                                res = np.float32(math.sqrt(uX * uX + uY * uY + uZ * uZ) / diag)
                                res = np.float32(np.exp(np.float32(-((res * res) * np.float32(10.0)))))
                                a = acc[nc]
                                a[0] = np.float32(a[0] + np.float32(rec[3]) * res)
                                a[1] = np.float32(a[1] + np.float32(rec[4]) * res)
                                a[2] = np.float32(a[2] + np.float32(rec[5]) * res)
                                a[3] = np.float32(a[3] + res)
            color = acc
        result = []
        for cc in sorted(occ, key=lambda c: c[0] + c[1] * g + c[2] * g * g):
            rank, idx, rec = min(occ[cc])
            q = quant(pts[idx])
            if mode == "none":
                rgb = cols[idx]
            elif mode == "weighted":
                a = color[cc]
                rgb = tuple(int(np.float32(a[j] / a[3])) for j in range(3))
            else:
                rgb = decode(color[cc])
            result.append((idx, (*q, *rgb)))
        subs[pid] = result
        point_count[pid] = len(result)

    for level in range(1, depth + 1):
        for pid in range(M):
            if keys[pid][0] == level and internal[pid]:
                subsample(pid)

    cursor = leaf_offset

    def place(pid):
        nonlocal cursor
        for c in children[pid]:
            if c != -1:
                place(c)
        if internal[pid]:
            data_idx[pid] = cursor
            cursor += point_count[pid]

    place(root)
    used = cursor
    if used <= E:
        for pid in range(M):
            if internal[pid]:
                for j, (idx, rec) in enumerate(subs[pid]):
                    lut[data_idx[pid] + j] = idx
                    out[data_idx[pid] + j] = rec
    return dict(keys=keys, dense_keys=[_dense_key(G, k) for k in keys], count=[counts[k] for k in keys],
                point_count=point_count, data_idx=data_idx, children=children, finished=finished,
                internal=internal, lut=lut, out=out, used=used, capacity=E)


def _dense_key(G, k):
    l, x, y, z = k
    off = sum((G >> j) ** 3 for j in range(l))
    g = G >> l
    return off + x + y * g + z * g * g
