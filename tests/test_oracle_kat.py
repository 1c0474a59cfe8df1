"""CPU oracle against the reference's own known answers and closed forms (no GPU needed).

Sources of each pin:
  * level tables and node statistics: OctreeApi/OctreeLibrary/Tests/test_octree.cu:44-217
  * output capacity: test_point_cloud_host.cu:46 (E = floor(1 * 2.2f) = 2)
  * the 128^3 cuboid: test_initial_point_counting_sparse.cpp:10-39, test_merging_sparse.cpp:10-157,
    test_point_distribution_sparse.cpp:10-93, test_subsampling_sparse.cpp:9-69 (cuboid generator
    tools.cu:7-23, :166-190)
  * packed colours: kernel_helpers.cuh:35-53 and rp_subsampling.cuh:123-129, checked against
    arbitrary-precision integer sums
  * explicit-random ranking: published SplitMix64 and MurmurHash3 test vectors
"""
import math

import numpy as np
import pytest

import oracle_binding as O
from octree_amd import workloads as W


def test_level_tables_chunking_grid_512():
    depth, grids, nodes, offs, dense = O.level_tables(512)
    assert depth == 9
    assert dense == 153_391_689
    assert list(grids) == [512, 256, 128, 64, 32, 16, 8, 4, 2, 1]
    assert list(nodes) == [g ** 3 for g in grids]
    assert list(offs) == [0, 134217728, 150994944, 153092096, 153354240, 153387008, 153391104, 153391616,
                          153391680, 153391688]


@pytest.mark.parametrize("grid", [1, 3, 100, 1024])
def test_unsupported_chunking_grids_have_depth_zero(grid):
    assert O.level_tables(grid)[0] == 0


def test_node_statistics_of_dummy_octree():
    """fillDummyOctree (test_octree.cu:4-42) and its expected statistics (:155-178)."""
    pc = np.array([100, 350, 0, 800, 1300, 0, 700, 250, 50], np.uint32)
    children = np.full((9, 8), -1, np.int32)
    children[8] = [0, 1, -1, 3, 4, -1, 6, 7]
    flags = np.array([1] * 8 + [3], np.uint8)
    st = O.node_statistics(pc, children, flags, 8)
    assert st["internal"] == 1 and st["leaves"] == 6
    assert st["min"] == 100 and st["max"] == 1300
    assert st["max_leaf_depth"] == 1
    assert st["mean"] == pytest.approx(583.3, rel=1e-3)


@pytest.mark.parametrize("n,factor,expected", [(1, 2.2, 2), (5_138_448, 2.2, 11_304_586),
                                               (25_836_417, 2.2, 56_840_116), (119_701_547, 2.2, 263_343_408)])
def test_output_capacity_is_a_float_product(n, factor, expected):
    assert O.output_capacity(n, factor) == expected
    assert int(np.float32(n) * np.float32(factor)) == expected


def test_map_point_to_grid_closed_forms():
    # cell = floor((p - min) / (size / G)) clamped to G - 1, linear x + y G + z G^2
    assert O.map_point_to_grid(0.0, 0.0, 0.0, 512, 100.0, [0, 0, 0]) == 0
    assert O.map_point_to_grid(100.0, 100.0, 100.0, 512, 100.0, [0, 0, 0]) == 511 + 511 * 512 + 511 * 512 * 512
    assert O.map_point_to_grid(50.0, 0.0, 0.0, 512, 100.0, [0, 0, 0]) == 256
    # a point just below the minimum truncates towards zero (int64 cast), as in the reference
    assert O.map_point_to_grid(-1e-9, 0.0, 0.0, 512, 100.0, [0, 0, 0]) == 0
    # cuboid: x + 0.5 over [0.5, 127.5] with G = 128 lands in cell x (x = 127 is clamped)
    for x in range(128):
        assert O.map_point_to_grid(x + 0.5, 0.5, 0.5, 128, 127.0, [0.5, 0.5, 0.5]) == x


def test_packed_colour_encoding_and_wrap():
    e = O.encode_colors(255, 128, 7)
    assert e == (255 << 46) | (128 << 28) | (7 << 10) | 1
    assert O.decode_colors(e) == (255, 128, 7)
    # 300 points of red 255: the red sum 76500 exceeds the 16-bit decode mask
    total = (300 * O.encode_colors(255, 0, 0)) % 2 ** 64
    assert O.decode_colors(total) == ((76500 & 0xFFFF) // 300, 0, 0)
    # 1024 contributions: the 10-bit count wraps to 0 and carries into blue; the reference
    # divides by zero there, the oracle decodes to 0
    total = (1024 * O.encode_colors(1, 1, 1)) % 2 ** 64
    assert O.decode_colors(total) == (0, 0, 0)
    total = (1025 * O.encode_colors(1, 1, 1)) % 2 ** 64
    assert O.decode_colors(total) == (1025, 1025, 1026)


def test_random_rank_uses_published_mixers():
    # SplitMix64 (seed 1234567) first two outputs, and MurmurHash3_x86_32("", seed=1) = 0x514E28B7
    s0 = 6457827717110365317 >> 32
    s1 = 3203168211198807973 >> 32
    assert O.random_rank(1234567, 0, s0 ^ 1) == 0x514E28B7
    assert O.random_rank(1234567, 1, s1 ^ 0xFFFFFFFF) == 0x81F16F39
    assert O.random_rank(1234567, 0, s0) == 0  # fmix32(0) = 0


def test_random_rank_is_a_bijection_per_node():
    ranks = {O.random_rank(5, 77, i) for i in range(20000)}
    assert len(ranks) == 20000


@pytest.fixture(scope="module")
def cuboid():
    rec, bb, scale = W.cuboid_cloud(128)
    spec = W.BuildSpec(128, 10000, 4.2, 128, intra=True)
    r = O.build(rec, 128 ** 3, 15, 0, bb, scale, spec)
    assert r.status == 0
    return rec, r


def _levels(r, grid):
    depth, grids, nodes, offs, dense = O.level_tables(grid)
    return np.searchsorted(offs, r.key, side="right") - 1


def test_cuboid_counting(cuboid):
    rec, r = cuboid
    lvl = _levels(r, 128)
    assert (lvl == 0).sum() == 128 ** 3
    assert (r.count[lvl == 0] == 1).all()
    assert int(r.count[lvl == 0].sum()) == 128 ** 3


def test_cuboid_merging(cuboid):
    rec, r = cuboid
    assert len(r.key) == 2_396_745
    # root: dense index D - 1 and sparse id M - 1 (test_merging_sparse.cpp:148-156)
    assert r.count[-1] == 2_097_152 and r.key[-1] == 2_396_744
    lvl = _levels(r, 128)
    expected_counts = [1, 8, 64, 512, 4096, 32768, 262144, 2097152]
    for level, cnt in enumerate(expected_counts):
        m = lvl == level
        assert m.sum() == (128 >> level) ** 3
        assert (r.count[m] == cnt).all()
        finished = r.finished[m]
        internal = r.internal[m]
        if level <= 3:
            assert not finished.any() and not internal.any()
        elif level == 4:  # the 8^3-grid nodes (4096 points) become finished leaves under internal parents
            assert finished.all() and not internal.any()
        else:
            assert finished.all() and internal.all()
    leaves = r.finished & ~r.internal
    assert int(r.point_count[leaves].sum()) == 128 ** 3


def test_cuboid_distribution_membership(cuboid):
    """Every leaf point lies in its leaf's voxel (test_point_distribution_sparse.cpp:36-46)."""
    rec, r = cuboid
    xyz = rec.reshape(-1, 15)[:, :12].copy().view(np.float32).reshape(-1, 3)
    lvl = _levels(r, 128)
    leaves = np.nonzero(r.finished & ~r.internal)[0]
    total = 0
    for leaf in leaves[::97]:
        level = lvl[leaf]
        g = 128 >> level
        lin = int(r.key[leaf]) - sum((128 >> k) ** 3 for k in range(level))
        cx, cy, cz = lin % g, (lin // g) % g, lin // (g * g)
        a, n = int(r.data_idx[leaf]), int(r.point_count[leaf])
        p = xyz[r.lut[a:a + n]]
        vox = 128.0 / g
        assert (p[:, 0] > cx * vox).all() and (p[:, 0] < (cx + 1) * vox).all()
        assert (p[:, 1] > cy * vox).all() and (p[:, 1] < (cy + 1) * vox).all()
        assert (p[:, 2] > cz * vox).all() and (p[:, 2] < (cz + 1) * vox).all()
        total += n
    assert total > 0


def test_cuboid_subsampling(cuboid):
    """73 internal nodes; levels 7, 6, 5 keep 128^3, 64^3, 32^3 points (test_subsampling_sparse.cpp)."""
    rec, r = cuboid
    lvl = _levels(r, 128)
    assert r.internal.sum() == 64 + 8 + 1
    for level, keep in ((7, 128 ** 3), (6, 64 ** 3), (5, 32 ** 3)):
        m = r.internal & (lvl == level)
        assert (r.point_count[m] == keep).all(), level
    # intra-cell averaging of one-point cells passes the colour through: r = x mod 256
    used = r.used
    assert used == 4 * 128 ** 3
    assert (r.out["r"][:used] == (r.lut[:used] % 128)).all()


def test_cuboid_overflow_is_an_error():
    """N + I = 4N exceeds E = 2.2N: the reference writes past its buffers (A14)."""
    rec, bb, scale = W.cuboid_cloud(64)
    spec = W.BuildSpec(64, 2000, 2.2, 64)
    r = O.build(rec, 64 ** 3, 15, 0, bb, scale, spec, want_data=False)
    assert r.status == 3 and r.used > r.capacity


@pytest.mark.parametrize("spec_kw", [dict(chunking_grid=100), dict(chunking_grid=1024),
                                     dict(intra=True, weighted=True), dict(output_factor=0.5)])
def test_invalid_configurations_are_rejected(spec_kw):
    cloud = W.generate(W.CloudSpec(1000, 0, (10.0, 10.0, 4.0), 2, 1))
    _, cubic = W.bounding_box(cloud, 1000, 15, 0)
    spec = W.BuildSpec(**{**dict(chunking_grid=64, merging_threshold=100, subsampling_grid=32), **spec_kw})
    assert O.build(cloud, 1000, 15, 0, cubic, [0.001] * 3, spec).status == 1


def test_root_forced_finished_when_below_threshold():
    cloud = W.generate(W.CloudSpec(300, 1, (10.0, 10.0, 4.0), 2, 1))
    _, cubic = W.bounding_box(cloud, 300, 27, 1)
    r = O.build(cloud, 300, 27, 1, cubic, [0.01] * 3, W.BuildSpec(64, 1000, 2.2, 32))
    assert r.status == 0
    assert r.finished[-1] and not r.internal[-1] and r.point_count[-1] == 300 and r.data_idx[-1] == 0
    assert r.finished.sum() == 1 and r.used == 300
    assert r.hierarchy == bytes([1, 0]) + (300).to_bytes(4, "little") + bytes(8) + (300 * 18).to_bytes(8, "little")


def test_explicit_random_selection_is_uniform():
    """Each point of a cell is the winner with equal probability (chi-square over many nodes)."""
    rng_counts = np.zeros(4, np.int64)
    for node in range(4000):
        ranks = [O.random_rank(99, node, i) for i in range(4)]
        rng_counts[int(np.argmin(ranks))] += 1
    expected = rng_counts.sum() / 4
    chi2 = float(((rng_counts - expected) ** 2 / expected).sum())
    assert chi2 < 16.27  # p = 0.001 for 3 degrees of freedom
