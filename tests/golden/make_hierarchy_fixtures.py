"""Writes hierarchy_fixtures.json: what the reference's own Potree reader decodes from the
oracle's Potree export of three seeded trees.

The reference writes hierarchy.bin breadth-first over the finished children of each node in slot
order, with child-mask bit i for child slot i (potree_exporter.cu:62-111). Its own reader,
tools/potree_to_ply.py:50-116, walks the file breadth-first using the popcount of each child
mask and writes one PLY file per tree level plus one of all leaves. This script exports each
oracle tree (hierarchy.bin, octree.bin, metadata.json), runs that reader on the export with
runpy (in this container only: the reader lives under /root/reference and is not copied), and
keeps, per decoded level and for the leaves, the number of records and the wrapping sum of a
64-bit hash of every 18-byte record (order-free, so a build that orders records differently
inside a node decodes to the same values).

The tests read only the JSON: test_hierarchy_reader.py checks that the oracle's node levels
(from the node keys) and its current export decode to these groups; the GPU test does the same
for a build through the C ABI.

    python tests/golden/make_hierarchy_fixtures.py
"""
import contextlib
import io
import json
import os
import runpy
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "octree-cuda_amd"))

import oracle_binding as O  # noqa: E402
import potree_files as P  # noqa: E402
from octree_amd import workloads as W  # noqa: E402

READER = "/root/reference/tools/potree_to_ply.py"


def run_reference_reader(directory):
    """tools/potree_to_ply.py -i directory -o directory; returns [(points, record bytes)] per
    level in the reader's order, the leaf group last."""
    argv = sys.argv
    sys.argv = ["potree_to_ply.py", "-i", directory, "-o", directory]
    try:
        with contextlib.redirect_stdout(io.StringIO()):
            runpy.run_path(READER, run_name="__main__")
    finally:
        sys.argv = argv
    groups = []
    level = 0
    while os.path.exists(os.path.join(directory, f"converted_level_{level}.ply")):
        groups.append(P.read_ply(os.path.join(directory, f"converted_level_{level}.ply")))
        level += 1
    groups.append(P.read_ply(os.path.join(directory, "converted_level_leaf.ply")))
    return groups


def main():
    fixtures = {}
    for name in sorted(P.HIERARCHY_CASES):
        spec, build = P.hierarchy_case(name)
        cloud = W.generate(spec, threads=1)
        _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type, threads=1)
        r = O.build(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build, threads=1)
        assert r.status == 0, (name, r.status)
        with tempfile.TemporaryDirectory() as d:
            P.write_potree(d, r.hierarchy, r.out, cubic[3] - cubic[0], spec.scale, build.subsampling_grid)
            groups = run_reference_reader(d)
        levels = [P.group_digest(records) for _, records in groups[:-1]]
        leaves = P.group_digest(groups[-1][1])
        for (points, records), digest in zip(groups, levels + [leaves]):
            assert points == digest["points"], (name, points, digest)
        masks = np.frombuffer(r.hierarchy, O.HIER_DTYPE)["mask"]
        assert len(levels) >= 4, (name, "fewer than 4 exported levels")
        assert ((masks != 0) & (masks != 0xFF)).any(), (name, "no partial child mask")
        fixtures[name] = {"nodes": len(masks), "levels": levels, "leaves": leaves}
        print(name, spec.points, "points,", len(masks), "exported nodes,", len(levels), "levels,",
              [g["points"] for g in levels], "leaf points", leaves["points"])
    with open(os.path.join(HERE, "hierarchy_fixtures.json"), "w") as f:
        json.dump(fixtures, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
