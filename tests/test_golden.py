"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py from the oracle):
the generator reproduces each stored cloud, the oracle reproduces each stored result, and on the
GPU the C ABI build matches it (nodes, used size and hierarchy.bin exactly; per-node point sets,
positions and colours exactly, except the colours of the distance-weighted filter's internal
nodes, which the per-level test in test_gpu_parity.py bounds)."""
import glob
import os
import sys

import numpy as np
import pytest

import oracle_binding as O
from octree_amd import workloads as W
from parity import fetch, gpu_build, hierarchy_bytes, node_of_position, records_by_node

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_fixtures as MF  # noqa: E402

NAMES = sorted(MF.CASES)


def load(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def check(name, fx, summ, key, count, flags, parent, children, point_count, data_idx, lut, out, hierarchy,
          exact_internal_colors):
    used = int(fx["used"])
    assert summ == used, (name, summ, used)
    for label, got, want in (("key", key, fx["key"]), ("count", count, fx["count"]), ("flags", flags, fx["flags"]),
                             ("parent", parent, fx["parent"]), ("children", children, fx["children"]),
                             ("point_count", point_count, fx["point_count"])):
        np.testing.assert_array_equal(got, want, f"{name}: {label}")
    fin = (fx["flags"] & 1).astype(bool)
    np.testing.assert_array_equal(data_idx[fin], fx["data_idx"][fin], f"{name}: data offsets")
    assert hierarchy == fx["hierarchy"].tobytes(), f"{name}: hierarchy.bin"
    owner = node_of_position(fx["data_idx"], fx["point_count"], fin, used)
    go, gi, gc = records_by_node(owner, lut[:used], out[:used])
    fo, fi, fc = records_by_node(owner, fx["lut"], fx["out"])
    np.testing.assert_array_equal(gi, fi, f"{name}: per-node point sets")
    np.testing.assert_array_equal(gc[:, :3], fc[:, :3], f"{name}: positions")
    if exact_internal_colors:
        np.testing.assert_array_equal(gc[:, 3:], fc[:, 3:], f"{name}: colours")
    else:
        leaf_record = ~(fx["flags"] & 2).astype(bool)[go]
        np.testing.assert_array_equal(gc[leaf_record, 3:], fc[leaf_record, 3:], f"{name}: leaf colours")


@pytest.mark.parametrize("name", NAMES)
def test_generator_reproduces_fixture_cloud(name):
    fx = load(name)
    spec, _ = MF.case_spec(name)
    np.testing.assert_array_equal(W.generate(spec), fx["cloud"])
    _, cubic = W.bounding_box(fx["cloud"], spec.points, spec.stride, spec.cloud_type)
    assert cubic == list(fx["cubic"])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    fx = load(name)
    spec, build = MF.case_spec(name)
    r = O.build(fx["cloud"], spec.points, spec.stride, spec.cloud_type, list(fx["cubic"]), [spec.scale] * 3, build)
    assert r.status == 0
    check(name, fx, r.used, r.key, r.count, r.flags, r.parent, r.children, r.point_count, r.data_idx, r.lut, r.out,
          r.hierarchy, exact_internal_colors=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(name):
    print(f"case golden {name}", flush=True)
    fx = load(name)
    spec, build = MF.case_spec(name)
    s = gpu_build(fx["cloud"], spec.points, spec.stride, spec.cloud_type, list(fx["cubic"]), [spec.scale] * 3, build)
    try:
        summ, nodes, lut, out = fetch(s)
        check(name, fx, summ["used"], nodes.key, nodes.count, nodes.flags, nodes.parent, nodes.children,
              nodes.point_count, nodes.data_idx, lut, out, hierarchy_bytes(s),
              exact_internal_colors=not build.weighted)
    finally:
        s.close()
