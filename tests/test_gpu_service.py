"""The build service (continuous batching) on the GPU: every request's result equals the same
request built alone, and both equal the oracle's; the memory budget is respected, and failures
stay with their request.

Equality: node keys, counts, flags, links, point counts, data offsets, used size and per-node
point sets and positions exactly; colours exactly, except the distance-weighted filter, whose
float sums depend on the order points arrive in (+-1 per channel, as against the oracle).
"""
import numpy as np
import pytest

import oracle_binding as O
from octree_amd import api
from octree_amd import workloads as W
from parity import WEIGHTED_DIFF_FRACTION, check_against_oracle, fetch, node_of_position, records_by_node, synthetic

pytestmark = pytest.mark.gpu

MODES = [(False, False, False, False), (True, False, False, True), (False, True, False, False),
         (False, True, True, False), (False, True, True, True), (True, False, False, False)]


def _requests():
    out = []
    sizes = [20_000, 150_000, 45_000, 400_000, 80_000, 260_000]
    for i, (n, mode) in enumerate(zip(sizes, MODES)):
        cloud_type = i % 2
        cloud, spec, cubic = synthetic(n, cloud_type, box=(10.0 + 7 * i, 12.0, 5.0), clusters=3, seed=30 + i,
                                       offset=0.0 if cloud_type == 0 else 2e5)
        intra, inter, weighted, rnd = mode
        build = W.BuildSpec(64 if i % 3 else 128, 2000 + 500 * i, 3.0, 32 if i % 2 else 64, intra=intra,
                            inter=inter, weighted=weighted, random=rnd, seed=100 + i)
        out.append((cloud, spec, cubic, build))
    return out


def _session(req):
    cloud, spec, cubic, build = req
    s = api.Session(0)
    s.set_cloud(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3)
    s.configure(build)
    return s


def _same(a, b, weighted):
    summ_a, nodes_a, lut_a, out_a = a
    summ_b, nodes_b, lut_b, out_b = b
    assert summ_a == summ_b
    for f in ("key", "count", "flags", "parent", "children", "point_count", "data_idx"):
        np.testing.assert_array_equal(getattr(nodes_a, f), getattr(nodes_b, f), f)
    fin = nodes_a.finished
    used = summ_a["used"]
    owner = node_of_position(nodes_a.data_idx, nodes_a.point_count, fin, used)
    ga, ia, ca = records_by_node(owner, lut_a, out_a)
    gb, ib, cb = records_by_node(owner, lut_b, out_b)
    np.testing.assert_array_equal(ga, gb)
    np.testing.assert_array_equal(ia, ib, "per-node point sets")
    np.testing.assert_array_equal(ca[:, :3], cb[:, :3], "positions")
    d = np.abs(ca[:, 3:] - cb[:, 3:]).max(initial=0)
    assert d <= (1 if weighted else 0), f"colours differ by {d}"


def _service(kind, slots, budget):
    """kind "slots": the slot service; "batch": the batch service with `slots` workers."""
    return api.Service(0, slots=slots, memory_budget=budget, batch_size=None if kind == "slots" else 16)


@pytest.mark.parametrize("kind,slots,copies", [("slots", 3, 1), ("slots", 16, 8), ("batch", 2, 1), ("batch", 3, 8)])
def test_service_results_equal_builds_alone(kind, slots, copies):
    """copies x the 6 requests through `slots` slots (16 slots and 48 requests: every slot runs its
    two-deep pipeline of build and result copies several times) or batch workers."""
    reqs = _requests()
    alone = []
    differing = compared = 0
    for cloud, spec, cubic, build in reqs:
        s = _session((cloud, spec, cubic, build))
        s.init()
        s.generate()
        alone.append(fetch(s))
        s.close()
        weighted = bool(build.weighted and build.inter and not build.intra)
        orc = O.build(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build, keep=weighted)
        try:
            _, d, c = check_against_oracle(alone[-1], orc, weighted, sample_nodes=10**9, max_fraction=1.0)
        finally:
            if orc.handle is not None:
                orc.handle.close()
        differing += d
        compared += c
    assert differing <= WEIGHTED_DIFF_FRACTION * max(compared, 1), (differing, compared)
    mix = [i for _ in range(copies) for i in range(len(reqs))]
    print("case", kind, "service mix of", len(mix), "requests,", slots, "workers", flush=True)
    with _service(kind, slots, 64 << 30) as svc:
        sessions = [_session(reqs[i]) for i in mix]
        for s in sessions:
            svc.submit(s)
        for s in sessions:
            svc.wait(s)
        st = svc.stats()
        assert st["completed"] == len(mix) and st["failed"] == 0
        assert st["p99_ms"] >= st["p50_ms"] > 0
        for i, s in zip(mix, sessions):
            _same(fetch(s), alone[i], weighted=reqs[i][3].weighted)
            t = s.request_timing()
            assert t["total_ms"] >= t["run_ms"] > 0
            s.close()


@pytest.mark.parametrize("kind", ["slots", "batch"])
def test_service_keeps_the_memory_budget(kind):
    reqs = _requests()[:4]
    sessions = [_session(r) for r in reqs]
    # room for about one request at a time
    need = max(s.footprint() if kind == "slots" else api.batch_footprint([s]) for s in sessions)
    budget = int(need * 1.3)
    print("case", kind, "service budget", budget, flush=True)
    with _service(kind, 3, budget) as svc:
        for s in sessions:
            svc.submit(s)
        for s in sessions:
            svc.wait(s)
        st = svc.stats()
        assert st["completed"] == len(reqs)
        assert 0 < st["peak_reserved_bytes"] <= budget
    for s in sessions:
        s.close()


@pytest.mark.parametrize("kind", ["slots", "batch"])
def test_service_rejects_requests_above_the_budget(kind):
    req = _requests()[0]
    s = _session(req)
    with _service(kind, 2, 1 << 20) as svc:
        with pytest.raises(api.OctreeError, match="budget"):
            svc.submit(s)
    s.close()


@pytest.mark.parametrize("kind", ["slots", "batch"])
def test_service_failure_stays_with_its_request(kind):
    reqs = _requests()[:3]
    cloud, spec, cubic, build = reqs[1]
    import dataclasses
    bad = (cloud, spec, cubic, dataclasses.replace(build, output_factor=1.0))  # N + I > E
    sessions = [_session(reqs[0]), _session(bad), _session(reqs[2])]
    with _service(kind, 2, 64 << 30) as svc:
        for s in sessions:
            svc.submit(s)
        svc.wait(sessions[0])
        with pytest.raises(api.OctreeError, match="output"):
            svc.wait(sessions[1])
        svc.wait(sessions[2])
        st = svc.stats()
        assert st["completed"] == 2 and st["failed"] == 1
    for s in sessions:
        s.close()
