"""GPU builder (libOctreeApi.so through the C ABI) against the CPU oracle on seeded inputs.

Every case compares node keys, counts, flags, links, point counts, data offsets, the used size,
hierarchy.bin and the per-node point sets bit-exactly; colours are bit-exact except for the
distance-weighted filter (float sums in a different order and a different expf: +-1 per channel).
"""
import itertools
import json
import re

import numpy as np
import pytest

import oracle_binding as O
from octree_amd import workloads as W
from octree_amd import api
from parity import compare, fetch, gpu_build, run_case, synthetic

pytestmark = pytest.mark.gpu

MODES = [  # (intra, inter, weighted, random)
    (False, False, False, False),
    (True, False, False, False),
    (False, True, False, False),
    (False, True, True, False),
    (False, False, False, True),
    (True, False, False, True),
    (False, True, False, True),
    (False, True, True, True),
]


def _spec(G, g, T, mode, F=3.0, seed=7):
    intra, inter, weighted, rnd = mode
    return W.BuildSpec(G, T, F, g, intra=intra, inter=inter, weighted=weighted, random=rnd, seed=seed)


@pytest.mark.parametrize("cloud_type", [0, 1])
@pytest.mark.parametrize("mode", MODES, ids=lambda m: "-".join(n for n, f in zip(("intra", "inter", "weighted", "random"), m) if f) or "plain")
def test_small_cloud_all_modes(cloud_type, mode):
    print(f"case small cloud_type={cloud_type} mode={mode}", flush=True)
    cloud, spec, cubic = synthetic(60_000, cloud_type, offset=0.0 if cloud_type == 0 else 5e5)
    run_case(cloud, spec.points, spec.stride, cloud_type, cubic, [spec.scale] * 3, _spec(64, 32, 2000, mode))


def test_cuboid_known_answers():
    """test_merging_sparse.cpp / test_subsampling_sparse.cpp on the 128^3 cuboid (15-byte variant)."""
    print("case cuboid 128^3", flush=True)
    rec, bb, scale = W.cuboid_cloud(128)
    spec = W.BuildSpec(128, 10000, 4.2, 128, intra=True)
    g, orc = run_case(rec, 128 ** 3, 15, 0, bb, scale, spec)
    summ, nodes, lut, out = g
    assert summ["sparse_nodes"] == 2_396_745
    assert nodes.count[-1] == 128 ** 3
    assert summ["internal_nodes"] == 73
    # intra-cell averaging of single-point cells passes colours through
    assert (out["r"][:128 ** 3] == (lut[:128 ** 3] % 128)).all()


@pytest.mark.parametrize("G,g,T", [(512, 128, 10000), (256, 64, 5000), (32, 16, 500)])
def test_grids_and_thresholds(G, g, T):
    print(f"case grids G={G} g={g} T={T}", flush=True)
    cloud, spec, cubic = synthetic(300_000, 0, box=(100.0, 100.0, 20.0), clusters=8, seed=3)
    run_case(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, _spec(G, g, T, (False, True, False, False)))


def test_root_as_single_leaf():
    """N < threshold: the reference walks off the tree (A13); the root becomes the only leaf."""
    cloud, spec, cubic = synthetic(500, 0)
    g, orc = run_case(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, _spec(64, 32, 10000, MODES[0]))
    summ = g[0]
    assert summ["internal_nodes"] == 0 and summ["leaf_nodes"] == 1 and summ["used"] == 500


@pytest.mark.parametrize("n", [1, 2, 7])
def test_tiny_clouds(n):
    cloud, spec, cubic = synthetic(max(n, 2), 1)
    cloud = cloud[: n * spec.stride]
    _, cubic = W.bounding_box(cloud, n, spec.stride, 1)
    if cubic[3] - cubic[0] <= 0:
        cubic = [cubic[0], cubic[1], cubic[2], cubic[0] + 1.0, cubic[1] + 1.0, cubic[2] + 1.0]
    run_case(cloud, n, spec.stride, 1, cubic, [spec.scale] * 3, _spec(8, 4, 0, MODES[2], F=40.0))


def test_threshold_zero_every_node_internal():
    cloud, spec, cubic = synthetic(5_000, 0)
    run_case(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, _spec(16, 8, 0, MODES[1], F=16.0))


def test_c1_shape_reduced():
    """The coin-like workload's generator and processing at 1/5 of its size."""
    wl = W.C1
    spec = W.CloudSpec(1_000_000, 0, wl.cloud.box, wl.cloud.clusters, wl.cloud.seed)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, 0)
    run_case(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, wl.build)


def test_split_buckets():
    """Dense clusters in a large box: two coarse buckets hold ~233K points each, above the fine
    sort's single-workgroup limit, so their histograms, dense counts and permutation come from
    the chunked path."""
    print("case split buckets n=3M G=512 g=128", flush=True)
    cloud, spec, cubic = synthetic(3_000_000, 0, box=(1000.0, 1000.0, 100.0), clusters=4, seed=5)
    run_case(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3,
             _spec(512, 128, 10000, (False, True, False, False)))


def test_statistics_report_matches_oracle(tmp_path):
    """statistics.json of a GPU build (ocpi_export_json_report; updateNodeStatistics,
    octree.cu:79-157) against the oracle's node statistics of its own tree, which the reference's
    fillDummyOctree known answers pin: node and leaf counts, min / max points per leaf and the
    deepest leaf exactly, the float mean and standard deviation to float rounding; and the point
    distribution histogram (ocpi_export_distribution_histogram) against the leaf counts."""
    print("case statistics report 300000 f32 G=256 g=64 T=5000", flush=True)
    cloud, spec, cubic = synthetic(300_000, 0, box=(100.0, 100.0, 20.0), clusters=8, seed=17)
    build = _spec(256, 64, 5000, MODES[2])
    orc = O.build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
    assert orc.status == 0
    want = O.node_statistics(orc.point_count, orc.children, orc.flags, len(orc.key) - 1)
    s = gpu_build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
    try:
        path = tmp_path / "statistics.json"
        s.export_json_report(str(path))
        got = json.loads(path.read_text())
        hist = {}
        for width in (500, 0):  # 0: the reference's automatic width (histogram.cu:39-119)
            h = tmp_path / f"hist{width}.html"
            s.export_distribution_histogram(str(h), width)
            text = h.read_text()
            data = text[text.index("data:[") + 6:]
            hist[width] = np.array([int(v) for v in data[:data.index("]")].split(",")], np.int64)
    finally:
        s.close()
    nodes, dist = got["resultNodes"], got["pointDistribution"]
    assert nodes["leafNodeAmount"] == want["leaves"]
    assert nodes["internalNodeAmount"] == want["internal"]
    assert nodes["octreeNodes"] == want["leaves"] + want["internal"]
    assert nodes["maxLeafDepth"] == want["max_leaf_depth"]
    assert dist["minPointsPerNode"] == want["min"] and dist["maxPointsPerNode"] == want["max"]
    assert dist["meanPointsPerLeafNode"] == pytest.approx(want["mean"], rel=1e-5)
    assert dist["stdevPointsPerLeafNode"] == pytest.approx(want["stdev"], rel=1e-5)
    assert got["overallNodes"]["sparseOctreeNodes"] == len(orc.key)
    # the distribution histogram: leaves binned by point count from the smallest leaf up
    leaf_counts = orc.point_count[orc.finished & ~orc.internal].astype(np.int64)
    want_hist = np.bincount((leaf_counts - leaf_counts.min()) // 500)
    np.testing.assert_array_equal(hist[500], want_hist, "histogram with bin width 500")
    assert hist[0].sum() == want["leaves"]


def test_memory_report_and_automatic_histogram_width(tmp_path):
    """memory_report.html and the memory figures of statistics.json are the library's device
    allocation counters: at the end of the build at least the session's resident buffers (the
    uploaded cloud, the two dense grids of D cells, LUT and output of E records) are allocated,
    and at most the peak. The automatic histogram width (bin width 0: Scott's rule, 3.5 stdev /
    leaves^(1/3) rounded up, histogram.cu:44-48) bins the oracle's leaf counts: the width in the
    label is that of the oracle's statistics and every bin holds the leaves of its range."""
    print("case memory report and automatic histogram width 300000 f32 G=256 g=64 T=5000", flush=True)
    cloud, spec, cubic = synthetic(300_000, 0, box=(100.0, 100.0, 20.0), clusters=8, seed=17)
    build = _spec(256, 64, 5000, MODES[2])
    orc = O.build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
    assert orc.status == 0
    want = O.node_statistics(orc.point_count, orc.children, orc.flags, len(orc.key) - 1)
    s = gpu_build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
    try:
        summ = s.summary()
        s.export_memory_report(str(tmp_path / "memory_report.html"))
        s.export_json_report(str(tmp_path / "statistics.json"))
        s.export_distribution_histogram(str(tmp_path / "hist.html"), 0)
    finally:
        s.close()
    text = (tmp_path / "memory_report.html").read_text()
    peak = float(re.search(r"Peak device allocation: ([0-9.eE+-]+) MB", text).group(1))
    now = float(re.search(r"Allocated at session end: ([0-9.eE+-]+) MB", text).group(1))
    mem = json.loads((tmp_path / "statistics.json").read_text())["memory"]
    assert peak == pytest.approx(mem["peak"] / 1e6, rel=1e-4)
    assert now == pytest.approx(mem["allocated"] / 1e6, rel=1e-4)
    resident = spec.points * spec.stride + 2 * 4 * summ["dense_nodes"] + summ["capacity"] * (4 + 18)
    assert mem["allocated"] >= resident, (mem["allocated"], resident)
    assert mem["allocated"] <= mem["peak"]

    html = (tmp_path / "hist.html").read_text()
    width = int(re.search(r"binWidth\((\d+)\)", html).group(1))
    scott = 3.5 * want["stdev"] / want["leaves"] ** (1.0 / 3.0)
    if abs(scott - round(scott)) > 1e-3:  # away from a rounding boundary the width is exact
        assert width == int(np.ceil(scott)), (width, scott)
    else:
        assert abs(width - scott) <= 1.0
    data = html[html.index("data:[") + 6:]
    got = np.array([int(v) for v in data[:data.index("]")].split(",")], np.int64)
    leaf_counts = orc.point_count[orc.finished & ~orc.internal].astype(np.int64)
    low = int(leaf_counts.min())
    expected = np.bincount((leaf_counts - low) // width, minlength=(int(leaf_counts.max()) - low) // width + 1)
    np.testing.assert_array_equal(got, expected, f"histogram with the automatic width {width}")


@pytest.mark.parametrize("G,g,T,n,mode,F", [
    (64, 1, 2000, 50_000, MODES[2], 3.0),        # one cell per node: every node keeps one point
    (64, 3, 2000, 50_000, MODES[3], 3.0),        # rows shorter than a bitmap word, pencils half empty
    (64, 37, 2000, 200_000, MODES[7], 3.0),      # odd: a leaf's octant of its parent grid is not exact
    (128, 100, 5000, 200_000, MODES[5], 3.0),    # not a power of two
    (256, 256, 5000, 300_000, MODES[2], 10.0),   # 2 MB occupancy bitmap per node
], ids=["g1", "g3", "g37", "g100", "g256"])
def test_uneven_and_large_subsampling_grids(G, g, T, n, mode, F):
    """Subsampling grids the reference accepts (any size) beyond the powers of two up to 128."""
    print(f"case subsampling grid G={G} g={g} T={T} n={n} mode={mode}", flush=True)
    cloud, spec, cubic = synthetic(n, 0, box=(100.0, 100.0, 20.0), clusters=8, seed=3)
    run_case(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, _spec(G, g, T, mode, F=F))


def test_level_in_several_node_batches():
    """A level whose occupancy bitmaps exceed the builder's 6 GiB bitmap budget is subsampled in
    batches of nodes. At g = 512 a node's bitmap and prefix take 32 MiB, so a batch holds 192
    nodes; level 2 of this tree has 321 internal nodes and takes two batches, the second starting
    at a node and a tile other than the level's first."""
    print("case node batches G=64 g=512 T=400 n=400000", flush=True)
    cloud, spec, cubic = synthetic(400_000, 0, box=(100.0, 100.0, 20.0), clusters=8, seed=3)
    build = _spec(64, 512, 400, MODES[6], F=12.0)
    _, orc = run_case(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build, threads=4)
    starts = np.cumsum([0] + [(64 >> l) ** 3 for l in range(7)])
    level = np.searchsorted(starts, orc.key, side="right") - 1
    assert np.bincount(level[orc.internal]).max() > 192, "no level needs two batches"


def test_points_outside_the_box_fail_the_build():
    """Points below the cubic box's minimum (the reference turns their negative grid coordinates
    into indices past its grids; points above the maximum are clamped into the top cells, as
    there) make ocpi_generate_octree return -1 with a message; the same session then builds the
    cloud with its real box, equal to the oracle's build."""
    print("case points outside the box n=20000", flush=True)
    cloud, spec, cubic = synthetic(20_000, 0, seed=21)
    side = cubic[3] - cubic[0]
    half = [cubic[0] + side / 2, cubic[1] + side / 2, cubic[2] + side / 2, cubic[3], cubic[4], cubic[5]]
    build = _spec(64, 32, 2000, MODES[2])
    s = api.Session(0)
    try:
        with pytest.raises(api.OctreeError, match="outside the cubic bounding box"):
            gpu_build(cloud, spec.points, spec.stride, 0, half, [spec.scale] * 3, build, sess=s)
        gpu_build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build, sess=s)
        orc = O.build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
        compare(fetch(s), orc)
    finally:
        s.close()


def test_rebuild_on_one_session_with_other_clouds():
    """Clouds of other shapes and precisions built one after the other on one session at chunking
    grid 512, where the fine sort writes the level-0 occupancy bits and only the level-0 counts that
    are not zero: cells that one cloud occupies and the next leaves empty keep the earlier counts
    (and leaf ids), which no pass of the next build may read. Each build equals the oracle's."""
    s = api.Session(0)
    try:
        for n, cloud_type, box, seed, mode in ((150_000, 0, (100.0, 100.0, 20.0), 31, MODES[2]),
                                                (120_000, 1, (50.0, 80.0, 60.0), 32, MODES[6]),
                                                (150_000, 0, (100.0, 100.0, 20.0), 33, MODES[5])):
            print(f"case rebuild G=512 g=128 T=1000 n={n} cloud_type={cloud_type} seed={seed} mode={mode}", flush=True)
            cloud, spec, cubic = synthetic(n, cloud_type, box=box, clusters=4, seed=seed)
            build = _spec(512, 128, 1000, mode, F=6.0)  # these clouds need N + I = 3.7 N
            gpu_build(cloud, spec.points, spec.stride, cloud_type, cubic, [spec.scale] * 3, build, sess=s)
            orc = O.build(cloud, spec.points, spec.stride, cloud_type, cubic, [spec.scale] * 3, build)
            compare(fetch(s), orc)
    finally:
        s.close()
