"""The C++ oracle against the pure-Python brute force (tests/bruteforce.py) on tiny clouds.

The brute force derives the tree from the point set directly and subsamples with dictionaries, so
agreement pins the oracle's dense-grid / K4-style formulation, its leaf block order, post-order
placement, winners and colour filters. Weighted colours use float32 in both, in the same order;
only expf differs (libm vs numpy), hence +-1.
"""
import numpy as np
import pytest

import bruteforce as BF
import oracle_binding as O
from octree_amd import workloads as W

MODES = [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (0, 1, 1, 0), (0, 0, 0, 1), (1, 0, 0, 1), (0, 1, 1, 1),
         (1, 1, 0, 0), (0, 0, 1, 0)]


def _compare(r, bf, tol):
    assert r.status == 0
    assert list(r.key) == bf["dense_keys"]
    assert list(r.count) == bf["count"]
    assert list(r.point_count) == bf["point_count"]
    assert list(r.finished) == bf["finished"]
    assert list(r.internal) == bf["internal"]
    assert r.children.tolist() == bf["children"]
    fin = r.finished
    assert [int(v) for v in r.data_idx[fin]] == [v for v, f in zip(bf["data_idx"], bf["finished"]) if f]
    assert r.used == bf["used"]
    assert list(r.lut[:r.used]) == bf["lut"][:bf["used"]]
    o = r.out[:r.used]
    got = np.stack([o[f].astype(np.int64) for f in ("x", "y", "z", "r", "g", "b")], 1)
    exp = np.array(bf["out"][:bf["used"]], dtype=np.int64)
    np.testing.assert_array_equal(got[:, :3], exp[:, :3])
    assert np.abs(got[:, 3:] - exp[:, 3:]).max(initial=0) <= tol


@pytest.mark.parametrize("cloud_type", [0, 1])
@pytest.mark.parametrize("mode", MODES)
def test_oracle_matches_bruteforce(cloud_type, mode):
    intra, inter, weighted, rnd = mode
    spec = W.CloudSpec(2500, cloud_type, (10.0, 10.0, 4.0), 3, 7 + cloud_type, offset=0.0 if cloud_type == 0 else 1e3)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, cloud_type)
    b = W.BuildSpec(8, 100, 2.2, 4, intra=bool(intra), inter=bool(inter), weighted=bool(weighted),
                    random=bool(rnd), seed=5)
    r = O.build(cloud, spec.points, spec.stride, cloud_type, cubic, [spec.scale] * 3, b)
    bf = BF.build(cloud.tobytes(), spec.points, spec.stride, cloud_type, cubic, [spec.scale] * 3, b)
    _compare(r, bf, tol=1 if (weighted and inter and not intra) else 0)


@pytest.mark.parametrize("G,g,T", [(2, 1, 0), (4, 3, 10), (16, 2, 50), (8, 5, 0)])
def test_oracle_matches_bruteforce_odd_grids(G, g, T):
    spec = W.CloudSpec(1200, 0, (5.0, 3.0, 2.0), 2, 3)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, 0)
    b = W.BuildSpec(G, T, 30.0, g, inter=True)
    r = O.build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, b)
    bf = BF.build(cloud.tobytes(), spec.points, spec.stride, 0, cubic, [spec.scale] * 3, b)
    _compare(r, bf, tol=0)


def test_oracle_thread_count_does_not_change_results():
    spec = W.CloudSpec(20000, 1, (20.0, 20.0, 5.0), 4, 9, offset=1e6)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, 1)
    b = W.BuildSpec(32, 500, 3.0, 16, inter=True, random=True, seed=3)
    r1 = O.build(cloud, spec.points, spec.stride, 1, cubic, [spec.scale] * 3, b, threads=1)
    r8 = O.build(cloud, spec.points, spec.stride, 1, cubic, [spec.scale] * 3, b, threads=8)
    np.testing.assert_array_equal(r1.lut, r8.lut)
    np.testing.assert_array_equal(r1.out, r8.out)
    assert r1.hierarchy == r8.hierarchy
