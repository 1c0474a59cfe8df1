"""GPU builds of the three single-cloud workloads at their full sizes against the CPU oracle (run
with up to 16 host threads): 5,138,448 x 15 B implicit + inter; 25,836,417 x 15 B explicit +
intra; 119,701,547 x 27 B f64 implicit + inter + distance-weighted, the benchmark workload.

Bit-exact: node keys, dense counts, flags, parents, children, point counts, data offsets of the
exported nodes (whose ranges must tile the used output), the used size, and every exported
node's records as a set, compared through order-free per-node sums of 64-bit record hashes
(point index, quantised position, colour). The subsampled colours of the distance-weighted
filter are left out of those sums and checked level by level (+-1 per channel against the
oracle fed with the GPU's own child records) on a seeded sample of internal nodes. hierarchy.bin
is a function of the compared node arrays; its writer is compared byte for byte with the
oracle's in the smaller parity cases (writing the 4.7 GB octree.bin here would only time the
disk).
"""
import os
import time

import numpy as np
import pytest

import oracle_binding as O
from octree_amd import workloads as W
from parity import check_against_oracle, fetch, gpu_build

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)
SAMPLE_NODES = 300  # internal nodes whose weighted colours are recomputed by the oracle


def _log(t0, msg):
    print(f"  [{time.time() - t0:6.1f} s] {msg}", flush=True)


def full_compare(cloud, n, stride, cloud_type, cubic, scale, build, t0):
    """GPU build against the oracle: node arrays, tiling of the output, record sets per node and,
    for the distance-weighted filter, sampled per-level colours."""
    weighted = bool(build.weighted and build.inter and not build.intra)
    orc = O.build(cloud, n, stride, cloud_type, cubic, scale, build, threads=THREADS, keep=weighted)
    assert orc.status == 0, f"oracle failed with status {orc.status}"
    _log(t0, f"oracle built with {THREADS} threads: {len(orc.key):,} nodes, {orc.used:,} records")
    s = gpu_build(cloud, n, stride, cloud_type, cubic, scale, build)
    try:
        g = fetch(s)
        summ, nodes, lut, out = g
        _log(t0, "GPU built, results copied")

        checked, _, _ = check_against_oracle(g, orc, weighted, sample_nodes=SAMPLE_NODES)
        _log(t0, f"record sets of {checked:,} nodes equal")
        return summ
    finally:
        s.close()
        if orc.handle is not None:
            orc.handle.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", ["C1", "C2", "C3"])
def test_full_size_workload(name):
    t0 = time.time()
    wl = W.WORKLOADS[name]
    spec, build = wl.cloud, wl.build
    print(f"case full size {name}: {wl.name}", flush=True)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type)
    _log(t0, "cloud generated")
    full_compare(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build, t0)


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


@pytest.fixture(scope="module", params=[1, 0], ids=["f64", "f32"])
def large_cloud(request):
    """4 M points in the benchmark cloud's shape, 27-byte f64 or 15-byte f32 records: the first
    levels take the multi-kernel path (device-wide scans, cleared compact arrays, node batches)
    rather than the one-workgroup setup."""
    ctype = request.param
    spec = W.CloudSpec(4_000_000, ctype, (400.0, 400.0, 40.0), 16, 9, offset=1e6 if ctype == 1 else 0.0)
    cloud = W.generate(spec)
    _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type)
    return cloud, spec, cubic


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", MODES, ids=lambda m: "-".join(n for n, f in zip(("intra", "inter", "weighted", "random"), m) if f) or "plain")
def test_large_cloud_every_mode(large_cloud, mode):
    t0 = time.time()
    cloud, spec, cubic = large_cloud
    intra, inter, weighted, rnd = mode
    build = W.BuildSpec(512, 10000, 3.0, 128, intra=intra, inter=inter, weighted=weighted, random=rnd, seed=13)
    print(f"case large cloud 4M type={spec.cloud_type} G=512 g=128 T=10000 mode={mode}", flush=True)
    summ = full_compare(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build, t0)
    assert summ["internal_nodes"] > 0
