"""The batched-build and serving workloads on the GPU against the CPU oracle (up to 16 host
threads), request by request.

Batched build: all 256 clouds of the seeded batch (sizes log-uniform in [10K, 2M] points, float32
and float64 records, boxes of 1 to 1000 units, grids 256 / 64, threshold 5000, implicit + inter)
submitted at once to the build service and built on overlapping slots.

Serving: a seeded request set from the serving loop's generator (sizes log-uniform in [50K, 20M]
points, float32 / float64 50/50, grids 512 / 128, threshold 10000, method and colour filter drawn
per request), the first two requests of each of the 8 (method x filter) modes, submitted at
Poisson arrival times (100 per second) so that they overlap on the device.

Every result: node arrays exactly, exported ranges tiling the output, every node's records as a
set (order-free hash sums), and for the distance-weighted filter the internal colours per level
(+-1) on a sample of nodes, with at most 1 % of all sampled channels one apart over the whole set
(src/main.cpp:55-109 is the per-request contract: one session per cloud).
"""
import collections
import os
import time

import numpy as np
import pytest

import oracle_binding as O
from octree_amd import api
from octree_amd import batch as B
from parity import WEIGHTED_DIFF_FRACTION, check_against_oracle, fetch

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)


def _weighted(build):
    return bool(build.weighted and build.inter and not build.intra)


def _check_all(requests, sessions, t0, sample_nodes):
    differing = compared = 0
    for i, (r, s) in enumerate(zip(requests, sessions)):
        weighted = _weighted(r.build)
        orc = O.build(r.cloud, r.points, r.stride, r.cloud_type, r.cubic, [r.scale] * 3, r.build, threads=THREADS,
                      keep=weighted)
        try:
            _, d, c = check_against_oracle(fetch(s), orc, weighted, sample_nodes=sample_nodes, max_fraction=1.0)
        except AssertionError as e:
            raise AssertionError(f"request {i} ({r.points} points, type {r.cloud_type}, {r.build}): {e}") from e
        finally:
            if orc.handle is not None:
                orc.handle.close()
        differing += d
        compared += c
        s.close()
        if (i + 1) % 32 == 0:
            print(f"  [{time.time() - t0:6.1f} s] {i + 1} of {len(requests)} results equal the oracle's", flush=True)
    print(f"  [{time.time() - t0:6.1f} s] {len(requests)} results equal the oracle's; weighted colours: "
          f"{differing} of {compared} sampled channels one apart", flush=True)
    assert differing <= WEIGHTED_DIFF_FRACTION * max(compared, 1)


@pytest.mark.timeout(900)
def test_batched_workload_against_oracle():
    t0 = time.time()
    print("case batched workload: 256 clouds through the build service, 8 slots", flush=True)
    reqs = B.batch_requests()
    sessions = [r.session() for r in reqs]
    with api.Service(0, slots=8, memory_budget=200 << 30) as svc:
        for s in sessions:
            svc.submit(s)
        for s in sessions:
            svc.wait(s)
        st = svc.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    print(f"  [{time.time() - t0:6.1f} s] built {sum(r.points for r in reqs):,} points", flush=True)
    _check_all(reqs, sessions, t0, sample_nodes=50)


@pytest.mark.timeout(900)
def test_batched_workload_batch_builder_against_oracle():
    """The same 256 clouds through the batched pipeline (one launch per stage for all clouds of a
    record type)."""
    t0 = time.time()
    print("case batched workload: 256 clouds through the batch builder", flush=True)
    reqs = B.batch_requests()
    sessions = [r.session() for r in reqs]
    with api.BatchBuilder(0) as bb:
        status = bb.build(sessions)
        t = bb.timings()
    assert (status == 0).all()
    print(f"  [{time.time() - t0:6.1f} s] built {sum(r.points for r in reqs):,} points: device {t['device_ms']:.1f} ms, "
          f"host {t['host_ms']:.1f} ms", flush=True)
    _check_all(reqs, sessions, t0, sample_nodes=50)


def serving_set(per_mode=2, rate=100.0, seed=B.SERVE_SEED):
    """The first `per_mode` requests of each mode from the serving generator's seeded stream,
    re-timed to Poisson arrivals at `rate` per second."""
    pool = B.serving_pool()
    stream = B.serving_requests(pool, rate=20.0, duration=60.0, seed=seed)
    taken = collections.Counter()
    picked = []
    for r in stream:
        mode = (r.build.intra, r.build.inter, r.build.weighted, r.build.random)
        if taken[mode] < per_mode:
            taken[mode] += 1
            picked.append(r)
    assert len(taken) == len(B.MODES) and min(taken.values()) == per_mode
    rng = np.random.default_rng(seed + 1)
    t = 0.0
    for r in picked:
        t += float(rng.exponential(1.0 / rate))
        r.arrival = t
    return picked


@pytest.mark.timeout(900)
@pytest.mark.parametrize("kind", ["slots", "batch"])
def test_serving_workload_against_oracle(kind):
    t0 = time.time()
    reqs = serving_set()
    print(f"case serving workload ({kind} service): {len(reqs)} requests, {sum(r.points for r in reqs):,} points, "
          f"arrivals over {reqs[-1].arrival:.2f} s", flush=True)
    sessions = [r.session() for r in reqs]
    with api.Service(0, slots=4 if kind == "slots" else 2, memory_budget=200 << 30,
                     batch_size=None if kind == "slots" else 32) as svc:
        start = time.perf_counter()
        for r, s in zip(reqs, sessions):
            delay = start + r.arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            svc.submit(s)
        for s in sessions:
            svc.wait(s)
        st = svc.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    print(f"  [{time.time() - t0:6.1f} s] served: p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms, "
          f"max queue depth {st['max_queue_depth']:.0f}", flush=True)
    _check_all(reqs, sessions, t0, sample_nodes=100)
