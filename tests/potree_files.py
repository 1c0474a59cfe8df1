"""Potree 2.0 export files as the tests read them: writing an oracle tree to a directory,
decoding hierarchy.bin breadth-first into per-level record groups, and the child-mask geometry.

Format (potree_exporter.cu:62-173, potree_exporter.cuh:28-35): hierarchy.bin holds one packed
22-byte entry {u8 type, u8 childMask, u32 points, u64 byteOffset, u64 byteSize} per exported
node, root first, then breadth-first; a node's children follow in the order of its set mask
bits; mask bit i stands for child slot i = 4 dx + 2 dy + dz (+x, +y, +z half of the parent cube,
point_count_propagation.cuh:62-70). octree.bin holds the 18-byte records {i32 x, y, z; u16 r, g,
b}, positions quantised against the cube's minimum (point_distributing.cuh:69-71).
"""
from __future__ import annotations

import json
import os

import numpy as np

from octree_amd import workloads as W

HIER_DTYPE = np.dtype([("type", "u1"), ("mask", "u1"), ("points", "<u4"), ("offset", "<u8"), ("size", "<u8")])
OUT_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("r", "<u2"), ("g", "<u2"), ("b", "<u2")])

# Three seeded trees with at least four exported levels and partial child masks.
# name: (points, cloud type, box, clusters, seed, offset, chunking grid, threshold, factor,
#        subsampling grid, intra, inter, random, rng seed)
HIERARCHY_CASES = {
    "f32_inter_implicit": (40_000, 0, (20.0, 20.0, 5.0), 4, 31, 0.0, 128, 500, 3.0, 16, False, True, False, 0),
    "f64_intra_explicit": (30_000, 1, (40.0, 10.0, 8.0), 3, 32, 1e5, 64, 300, 3.0, 16, True, False, True, 5),
    "f32_plain_explicit": (25_000, 0, (6.0, 6.0, 6.0), 6, 33, 0.0, 256, 200, 3.0, 32, False, False, True, 9),
}


def hierarchy_case(name):
    n, t, box, k, seed, off, G, T, F, g, intra, inter, rnd, rs = HIERARCHY_CASES[name]
    spec = W.CloudSpec(n, t, box, k, seed, offset=off)
    build = W.BuildSpec(G, T, F, g, intra=intra, inter=inter, random=rnd, seed=rs)
    return spec, build


def write_potree(directory, hierarchy: bytes, out: np.ndarray, side: float, scale: float, g: int):
    """hierarchy.bin, octree.bin (all records) and the metadata.json fields a reader uses."""
    entries = np.frombuffer(hierarchy, HIER_DTYPE)
    with open(os.path.join(directory, "hierarchy.bin"), "wb") as f:
        f.write(hierarchy)
    with open(os.path.join(directory, "octree.bin"), "wb") as f:
        f.write(np.ascontiguousarray(out).tobytes())
    meta = {"version": "2.0", "points": int(entries["points"].sum()),
            "hierarchy": {"firstChunkSize": 22 * len(entries), "stepSize": 100, "depth": 20},
            "scale": [scale] * 3, "spacing": side / g}
    with open(os.path.join(directory, "metadata.json"), "w") as f:
        json.dump(meta, f)


def read_ply(path):
    """(vertex count, record bytes) of a binary PLY file whose vertices are 18-byte records."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    count = None
    for line in data[:end].decode("ascii").splitlines():
        if line.startswith("element vertex "):
            count = int(line.split()[2])
    body = data[end:]
    assert count is not None and len(body) == 18 * count, (path, count, len(body))
    return count, body


def record_hash(records: bytes) -> np.ndarray:
    """64-bit hash of every 18-byte record: its three little-endian words (the last zero-padded),
    folded with the splitmix64 finaliser."""
    n = len(records) // 18
    raw = np.zeros((n, 24), np.uint8)
    raw[:, :18] = np.frombuffer(records, np.uint8).reshape(n, 18)
    w = raw.view("<u8").reshape(n, 3)

    def mix(z):
        z = z ^ (z >> np.uint64(30))
        z = z * np.uint64(0xBF58476D1CE4E5B9)
        z = z ^ (z >> np.uint64(27))
        z = z * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        return mix(w[:, 0] ^ mix(w[:, 1] ^ mix(w[:, 2] + np.uint64(0x9E3779B97F4A7C15))))


def group_digest(records: bytes) -> dict:
    """Record count and order-free wrapping sum of the record hashes of one group."""
    with np.errstate(over="ignore"):
        s = int(record_hash(records).sum(dtype=np.uint64)) if records else 0
    return {"points": len(records) // 18, "hash_sum": f"{s:016x}"}


def decode_bfs(hierarchy: bytes):
    """Breadth-first decode of hierarchy.bin: per entry (tree level, parent entry, child slot as
    given by the parent's mask bit). Children of an entry are the next unclaimed entries, one per
    set mask bit in ascending bit order. Raises if the entries do not form such a tree."""
    e = np.frombuffer(hierarchy, HIER_DTYPE)
    level = np.full(len(e), -1, np.int64)
    parent = np.full(len(e), -1, np.int64)
    slot = np.full(len(e), -1, np.int64)
    assert len(e) > 0
    level[0] = 0
    nxt = 1
    for i in range(len(e)):
        assert level[i] >= 0, f"entry {i} is nobody's child"
        mask = int(e["mask"][i])
        assert int(e["type"][i]) == (1 if mask == 0 else 0), f"entry {i}: type {e['type'][i]} with mask {mask}"
        for bit in range(8):
            if mask >> bit & 1:
                assert nxt < len(e), f"entry {i}: mask {mask:#x} names more children than the file holds"
                level[nxt] = level[i] + 1
                parent[nxt] = i
                slot[nxt] = bit
                nxt += 1
    assert nxt == len(e), f"{len(e) - nxt} entries are not reachable from the root"
    return e, level, parent, slot


def level_groups(hierarchy: bytes, octree: bytes):
    """Per tree level (root = 0) and for the leaves (mask 0): the concatenated records, in file
    order, as the breadth-first decode finds them."""
    e, level, _, _ = decode_bfs(hierarchy)
    levels = [[] for _ in range(int(level.max()) + 1)]
    leaves = []
    for i in range(len(e)):
        a, n = int(e["offset"][i]), int(e["size"][i])
        assert n == 18 * int(e["points"][i]), f"entry {i}: byteSize {n} for {e['points'][i]} points"
        assert a % 18 == 0 and a + n <= len(octree), f"entry {i}: records [{a}, {a + n}) outside octree.bin"
        rec = octree[a:a + n]
        levels[level[i]].append(rec)
        if int(e["mask"][i]) == 0:
            leaves.append(rec)
    return [b"".join(g) for g in levels], b"".join(leaves)


def digests(hierarchy: bytes, octree: bytes):
    levels, leaves = level_groups(hierarchy, octree)
    return [group_digest(g) for g in levels], group_digest(leaves)


def mask_geometry_violations(hierarchy: bytes, octree: bytes, side_units: float):
    """Entries whose records do not lie in the octant of the parent cube that their mask bit
    names. side_units = cube side in quantisation units (side / scale); the root cube is
    [0, side_units)^3. A child's octant is judged by the mean of its quantised positions, which
    lies inside the octant. Points on a dividing plane belong to the upper half (the cell mapping
    rounds down from there) but may quantise one unit below it, because 1 / scale of a float scale
    is not exact (point_distributing.cuh:69-71): a mean within two units of the plane is taken as
    consistent with either half."""
    e, level, parent, slot = decode_bfs(hierarchy)
    recs = np.frombuffer(octree, OUT_DTYPE)
    lo = np.zeros((len(e), 3))
    bad = []
    for i in range(1, len(e)):
        p = parent[i]
        half = side_units / 2.0 ** (level[i])
        bits = np.array([slot[i] >> 2 & 1, slot[i] >> 1 & 1, slot[i] & 1], np.float64)
        lo[i] = lo[p] + bits * half
        a, n = int(e["offset"][i]) // 18, int(e["points"][i])
        if n == 0:
            continue
        r = recs[a:a + n]
        mean = np.array([r["x"].mean(), r["y"].mean(), r["z"].mean()])
        centre = lo[p] + half
        for axis in range(3):
            if abs(mean[axis] - centre[axis]) <= 2.0:
                continue
            if (mean[axis] > centre[axis]) != bool(bits[axis]):
                bad.append(i)
                break
    return bad


def reserialise(hierarchy: bytes, order: str) -> bytes:
    """The same tree written the wrong way, for the tests that show the checks catch it:
    order="dfs" writes the entries depth-first (pre-order) instead of breadth-first;
    order="reversed_bits" keeps the order but puts child slot i at mask bit 7 - i."""
    e, level, parent, slot = decode_bfs(hierarchy)
    e = e.copy()
    if order == "reversed_bits":
        e["mask"] = np.array([int(f"{int(m):08b}"[::-1], 2) for m in e["mask"]], np.uint8)
        return e.tobytes()
    assert order == "dfs"
    children = [[] for _ in range(len(e))]
    for i in range(1, len(e)):
        children[parent[i]].append(i)
    seq, stack = [], [0]
    while stack:
        i = stack.pop()
        seq.append(i)
        stack.extend(reversed(children[i]))
    return e[seq].tobytes()
