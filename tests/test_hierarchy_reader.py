"""hierarchy.bin and octree.bin against what the reference's own Potree reader decodes.

tests/golden/hierarchy_fixtures.json holds, for three seeded trees, the per-level and leaf
record groups that tools/potree_to_ply.py:50-116 (the reference's reader: breadth-first, children
counted by the popcount of the child mask) produced from the oracle's export. Here:

* the oracle's node arrays, grouped by tree level from the node keys and without the writer, give
  the same groups (so the file the reader decoded was a breadth-first export of these nodes);
* the oracle's current export, decoded breadth-first by potree_files.decode_bfs, gives them too;
* every child's records lie in the octant of its parent that its mask bit names (bit i = slot
  4 dx + 2 dy + dz, potree_exporter.cu:99-111), which the reader itself does not look at;
* the same tree written depth-first, or with the mask bits reversed, fails these checks;
* on the GPU, a build through the C ABI exports files that decode to the same groups.
"""
import json
import os
import tempfile

import numpy as np
import pytest

import oracle_binding as O
import potree_files as P
from octree_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "hierarchy_fixtures.json")) as _f:
    FIXTURES = json.load(_f)
NAMES = sorted(P.HIERARCHY_CASES)


def oracle_case(name):
    spec, build = P.hierarchy_case(name)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type)
    r = O.build(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build)
    assert r.status == 0
    return spec, build, cubic, r


def groups_from_nodes(r, grid):
    """Per tree level and for the leaves: records of the finished nodes, from the node arrays."""
    depth, _, _, offs, _ = O.level_tables(grid)
    level = np.searchsorted(offs, r.key, side="right") - 1
    tree_level = depth - level
    fin = r.finished
    has_finished_child = np.array([any(c >= 0 and fin[c] for c in r.children[i]) for i in range(len(r.key))])
    out = r.out.tobytes()
    levels = [[] for _ in range(int(tree_level[fin].max()) + 1)]
    leaves = []
    for i in np.nonzero(fin)[0]:
        a, n = int(r.data_idx[i]), int(r.point_count[i])
        rec = out[18 * a:18 * (a + n)]
        levels[tree_level[i]].append(rec)
        if not has_finished_child[i]:
            leaves.append(rec)
    return [P.group_digest(b"".join(g)) for g in levels], P.group_digest(b"".join(leaves))


def side_units(cubic, scale):
    return (cubic[3] - cubic[0]) * (1.0 / scale)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_nodes_match_reference_reader(name):
    spec, build, cubic, r = oracle_case(name)
    levels, leaves = groups_from_nodes(r, build.chunking_grid)
    assert levels == FIXTURES[name]["levels"]
    assert leaves == FIXTURES[name]["leaves"]
    assert int(r.finished.sum()) == FIXTURES[name]["nodes"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_export_decodes_to_reference_reader_groups(name):
    spec, build, cubic, r = oracle_case(name)
    levels, leaves = P.digests(r.hierarchy, r.out.tobytes())
    assert levels == FIXTURES[name]["levels"]
    assert leaves == FIXTURES[name]["leaves"]
    assert P.mask_geometry_violations(r.hierarchy, r.out.tobytes(), side_units(cubic, spec.scale)) == []


@pytest.mark.parametrize("name", NAMES)
def test_depth_first_export_fails(name):
    spec, build, cubic, r = oracle_case(name)
    dfs = P.reserialise(r.hierarchy, "dfs")
    assert dfs != r.hierarchy
    try:
        levels, leaves = P.digests(dfs, r.out.tobytes())
    except AssertionError:
        return  # not even a breadth-first tree
    assert levels != FIXTURES[name]["levels"]


@pytest.mark.parametrize("name", NAMES)
def test_reversed_mask_bits_fail(name):
    spec, build, cubic, r = oracle_case(name)
    rev = P.reserialise(r.hierarchy, "reversed_bits")
    # the reference's reader only counts mask bits, so the groups are unchanged...
    assert P.digests(rev, r.out.tobytes()) == P.digests(r.hierarchy, r.out.tobytes())
    # ...but the children of every parent whose set of child slots is not symmetric under
    # s -> 7 - s no longer lie in the octants their bits name
    bad = P.mask_geometry_violations(rev, r.out.tobytes(), side_units(cubic, spec.scale))
    assert len(bad) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_export_decodes_to_reference_reader_groups(name):
    from parity import gpu_build

    print(f"case hierarchy reader {name}", flush=True)
    spec, build = P.hierarchy_case(name)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type)
    s = gpu_build(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build)
    try:
        with tempfile.TemporaryDirectory() as d:
            s.export_potree(d)
            with open(os.path.join(d, "hierarchy.bin"), "rb") as f:
                hierarchy = f.read()
            with open(os.path.join(d, "octree.bin"), "rb") as f:
                octree = f.read()
    finally:
        s.close()
    levels, leaves = P.digests(hierarchy, octree)
    assert levels == FIXTURES[name]["levels"]
    assert leaves == FIXTURES[name]["leaves"]
    assert P.mask_geometry_violations(hierarchy, octree, side_units(cubic, spec.scale)) == []
