"""Batched builds (ocpi_batch_build: one launch per stage for all clouds of a group) on the GPU.

Every session of a batch is compared with the oracle (node arrays exactly, exported ranges tiling
the output, every node's records as a set; distance-weighted colours per level within +-1 with at
most 1 % of the channels one apart over the batch), and its hierarchy.bin byte for byte. The
batches mix record types, chunking and subsampling grids, thresholds, all 8 (method x filter)
modes, and degenerate clouds (1, 2 and 7 points, a root-only tree, subsampling grids 1 and 3), so
that groups of one and of many clouds, clouds with and without subsampling and the large-bucket
path of the fine sort all run. Failures stay with their cloud.
"""
import dataclasses

import numpy as np
import pytest

import oracle_binding as O
from octree_amd import api
from octree_amd import workloads as W
from parity import WEIGHTED_DIFF_FRACTION, check_against_oracle, fetch, hierarchy_bytes, synthetic

pytestmark = pytest.mark.gpu

MODES = [(False, False, False, False), (True, False, False, False), (False, True, False, False),
         (False, True, True, False), (False, False, False, True), (True, False, False, True),
         (False, True, False, True), (False, True, True, True)]


def _req(n, ctype, seed, G, T, g, mode, box=(20.0, 14.0, 5.0), factor=3.0, clusters=3):
    cloud, spec, cubic = synthetic(max(n, 2), ctype, box=box, clusters=clusters, seed=seed,
                                   offset=0.0 if ctype == 0 else 3e5)
    if n < 2:
        cloud = cloud[: n * spec.stride]
        spec = dataclasses.replace(spec, points=n)
        _, cubic = W.bounding_box(cloud, n, spec.stride, ctype)
        cubic = cubic[:3] + [v + 1.0 for v in cubic[:3]]  # a single point: the CLI's box has no extent
    intra, inter, weighted, rnd = mode
    build = W.BuildSpec(G, T, factor, g, intra=intra, inter=inter, weighted=weighted, random=rnd, seed=seed * 7)
    return cloud, spec, cubic, build


def mixed_requests():
    reqs = []
    sizes = [30_000, 150_000, 4_000, 70_000, 260_000, 12_000, 45_000, 9_000]
    for i, mode in enumerate(MODES):
        for ctype in (0, 1):
            n = sizes[(i + 3 * ctype) % len(sizes)]
            reqs.append(_req(n, ctype, 100 + 2 * i + ctype, 64, 1000 + 300 * (i % 3), 32, mode))
    # another chunking grid, thresholds and subsampling grids in the same batch
    reqs.append(_req(120_000, 0, 140, 128, 2000, 64, MODES[2]))
    reqs.append(_req(120_000, 0, 141, 128, 500, 3, MODES[2]))
    reqs.append(_req(50_000, 1, 142, 128, 800, 1, MODES[0]))
    # a cloud whose densest bucket takes the split path of the fine sort (280K points in one
    # level-3 cell, above FINE_SPLIT = 196,608)
    spec = W.CloudSpec(600_000, 0, (100.0, 100.0, 100.0), 1, 143, heightfield_fraction=0.5, sigma=1.0)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type)
    reqs.append((cloud, spec, cubic, W.BuildSpec(64, 4000, 3.0, 32, inter=True, seed=5)))
    # degenerate clouds: tiny, and one whose root is its only leaf
    reqs.append(_req(1, 0, 150, 64, 1000, 32, MODES[2]))
    reqs.append(_req(2, 1, 151, 64, 1000, 32, MODES[3]))
    reqs.append(_req(7, 0, 152, 64, 1000, 32, MODES[0]))
    reqs.append(_req(500, 0, 153, 64, 1000, 32, MODES[2]))
    return reqs


def _session(req):
    cloud, spec, cubic, build = req
    s = api.Session(0)
    s.set_cloud(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3)
    s.configure(build)
    return s


def _check_batch(reqs, sessions, status):
    differing = compared = 0
    for i, (req, s) in enumerate(zip(reqs, sessions)):
        cloud, spec, cubic, build = req
        assert status[i] == 0, f"request {i} failed"
        weighted = bool(build.weighted and build.inter and not build.intra)
        orc = O.build(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build, keep=weighted)
        try:
            _, d, c = check_against_oracle(fetch(s), orc, weighted, sample_nodes=10**9, max_fraction=1.0)
            assert hierarchy_bytes(s) == orc.hierarchy, f"request {i}: hierarchy.bin differs"
        except AssertionError as e:
            raise AssertionError(f"request {i} ({spec.points} points, type {spec.cloud_type}, {build}): {e}") from e
        finally:
            if orc.handle is not None:
                orc.handle.close()
        differing += d
        compared += c
    assert differing <= WEIGHTED_DIFF_FRACTION * max(compared, 1), (differing, compared)


def test_batch_matches_the_oracle():
    reqs = mixed_requests()
    print(f"case batch of {len(reqs)} mixed requests", flush=True)
    sessions = [_session(r) for r in reqs]
    with api.BatchBuilder(0) as bb:
        status = bb.build(sessions)
        t = bb.timings()
    print(f"  device {t['device_ms']:.2f} ms, host {t['host_ms']:.2f} ms", flush=True)
    _check_batch(reqs, sessions, status)
    for s in sessions:
        s.close()


def test_batch_graph_replays_and_launch_shapes():
    """One builder (arena reused) builds the same batch six times: with plain launches, then three
    times through its captured graphs (the first build captures, the next two replay, with the
    host's table uploads, readbacks and checks between the replays), then with small persistent
    grids (every workgroup loops over many items) captured and plain. Every build equals the
    oracle's result."""
    reqs = [mixed_requests()[k] for k in (0, 3, 5, 9, 16, 19)]
    with api.BatchBuilder(0) as bb:
        for graphs, shape in ((False, (2048, 4096, 1024)), (True, (2048, 4096, 1024)), (True, (2048, 4096, 1024)),
                              (True, (2048, 4096, 1024)), (True, (7, 5, 3)), (False, (7, 5, 3))):
            print(f"case batch rebuild, graphs {graphs}, launch shape {shape}", flush=True)
            bb.use_graphs(graphs)
            bb.set_launch(*shape)
            sessions = [_session(r) for r in reqs]
            status = bb.build(sessions)
            _check_batch(reqs, sessions, status)
            for s in sessions:
                s.close()


def test_batch_failures_stay_with_their_request():
    good = mixed_requests()
    a, b = good[2], good[5]
    cloud, spec, cubic, build = good[8]
    overflow = (cloud, spec, cubic, dataclasses.replace(build, output_factor=1.0))  # N + I > E
    cloud2, spec2, cubic2, build2 = good[10]
    half = 0.5 * (cubic2[3] - cubic2[0])
    outside = (cloud2, spec2, [v + half for v in cubic2], build2)  # about half the points lie below the cube
    reqs = [a, overflow, b, outside]
    print("case batch with an overflowing and an outside-the-box request", flush=True)
    sessions = [_session(r) for r in reqs]
    with api.BatchBuilder(0) as bb:
        status = bb.build(sessions, raise_on_failure=False)
    assert list(status) == [0, -1, 0, -1]
    _check_batch([a, b], [sessions[0], sessions[2]], [0, 0])
    for s in sessions:
        s.close()
