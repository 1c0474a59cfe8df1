"""The command-line caller and the GPU bounding box on the MI355X: the box equals the reference
CLI's host formula exactly, and a CLI run writes the hierarchy the oracle predicts."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_binding as O
from hip_memory import DeviceBuffer
from octree_amd import api
from octree_amd import workloads as W
from parity import compare, fetch, gpu_build, synthetic

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "octree-cuda_amd", "bin", "octree-amd")


@pytest.mark.parametrize("cloud_type,n", [(0, 1), (0, 70_001), (1, 250_000)])
def test_gpu_bounding_box_equals_host_formula(cloud_type, n):
    print(f"case bounding box type={cloud_type} n={n}", flush=True)
    cloud, spec, _ = synthetic(n, cloud_type, box=(30.0, 11.0, 4.0), offset=0.0 if cloud_type == 0 else 3e5)
    real_h, cubic_h = W.bounding_box(cloud, n, spec.stride, cloud_type)
    buf = DeviceBuffer(cloud)
    try:
        real = (ctypes.c_double * 6)()
        cubic = (ctypes.c_double * 6)()
        rc = api.library().ocpi_compute_bounding_box(buf.ptr, n, spec.stride, cloud_type, 0, real, cubic)
        assert rc == 0, api.last_error()
        assert list(real) == real_h
        assert list(cubic) == cubic_h
    finally:
        buf.close()


def expected_metadata(hierarchy, side, scale, g, flags):
    """metadata.json as the reference's exporter writes it (potree_exporter.cu:8-32, 113-173): the
    points and nodes of the exported hierarchy (22-byte entries, u32 point count at byte 2), the
    cubic side, spacing = side / g, the scale as set, zero offset, in this key order."""
    nodes = len(hierarchy) // 22
    points = sum(int.from_bytes(hierarchy[22 * i + 2:22 * i + 6], "little") for i in range(nodes))
    return {
        "version": "2.0", "name": "GpuPotreeConverter", "description": "Generated with GpuPotreeConverter",
        "points": points, "projection": "", "flags": list(flags),
        "hierarchy": {"firstChunkSize": 22 * nodes, "stepSize": 100, "depth": 20},
        "offset": [0, 0, 0], "scale": [scale] * 3, "spacing": side / g,
        "boundingBox": {"min": [0, 0, 0], "max": [side] * 3}, "encoding": "default",
        "attributes": [
            {"name": "position", "description": "", "size": 12, "numElements": 3, "elementSize": 4,
             "type": "int32", "min": [0, 0, 0], "max": [side] * 3},
            {"name": "rgb", "description": "", "size": 6, "numElements": 3, "elementSize": 2, "type": "uint16",
             "min": [0, 0, 0], "max": [65536] * 3},
        ],
    }


def test_cli_writes_the_oracle_hierarchy(tmp_path):
    print("case cli 120000 points f64 inter", flush=True)
    cloud, spec, cubic = synthetic(120_000, 1, box=(25.0, 25.0, 6.0), offset=1e5)
    path = tmp_path / "cloud.bin"
    cloud.tofile(path)
    out = tmp_path / "potree"
    r = subprocess.run([CLI, "-f", str(path), "-o", str(out), "-p", str(spec.points), "-t", "double", "-d",
                        f"27,{spec.scale}", "-g", "64,32", "-m", "2000", "-e", "3.0", "-i"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    build = W.BuildSpec(64, 2000, 3.0, 32, inter=True)
    orc = O.build(cloud, spec.points, 27, 1, cubic, [float(np.float32(spec.scale))] * 3, build)
    assert orc.status == 0
    assert (out / "hierarchy.bin").read_bytes() == orc.hierarchy
    meta = json.loads((out / "metadata.json").read_text())
    # the CLI uses the replacement scheme, as the reference's (src/argparser.h:160)
    assert meta == expected_metadata(orc.hierarchy, cubic[3] - cubic[0], float(np.float32(spec.scale)), 32,
                                     ("REPLACING", "IMPLICIT-RANDOM-SUBSAMPLING", "INTER-CELL-COLOR-FILTERING"))
    assert list(meta) == list(expected_metadata(b"", 1.0, 1.0, 1, ())), "key order"
    assert os.path.getsize(out / "octree.bin") == int(np.float32(spec.points) * np.float32(3.0)) * 18
    for name in ("statistics.json", "point_distribution.html", "memory_report.html"):
        assert (out / name).exists(), name


@pytest.mark.parametrize("replacement,random,intra,inter,flags", [
    (False, True, True, False, ("ADDITIVE", "EXPLICIT-RANDOM-SUBSAMPLING", "INTRA-CELL-COLOR-FILTERING")),
    (True, False, False, False, ("REPLACING", "IMPLICIT-RANDOM-SUBSAMPLING")),
    (False, False, False, True, ("ADDITIVE", "IMPLICIT-RANDOM-SUBSAMPLING", "INTER-CELL-COLOR-FILTERING")),
])
def test_export_potree_metadata(tmp_path, replacement, random, intra, inter, flags):
    """ocpi_export_potree: hierarchy.bin equals the oracle's, octree.bin holds the built records and
    a zero tail, and metadata.json has the reference exporter's fields for each flag combination."""
    print(f"case export flags={flags}", flush=True)
    cloud, spec, cubic = synthetic(40_000, 0, seed=9)
    build = W.BuildSpec(64, 2000, 3.0, 32, intra=intra, inter=inter, random=random, replacement=replacement, seed=3)
    orc = O.build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
    assert orc.status == 0
    s = gpu_build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
    try:
        s.export_potree(str(tmp_path))
        g = fetch(s)
    finally:
        s.close()
    compare(g, orc)  # the build's nodes and per-node record sets equal the oracle's
    assert (tmp_path / "hierarchy.bin").read_bytes() == orc.hierarchy
    # octree.bin: all E records (potree_exporter.cu:56), the used ones as built, a zero tail (A15)
    summ, out = g[0], g[3]
    data = (tmp_path / "octree.bin").read_bytes()
    assert len(data) == orc.capacity * 18
    assert data[:summ["used"] * 18] == out.tobytes()
    assert not any(data[summ["used"] * 18:])
    meta = json.loads((tmp_path / "metadata.json").read_text())
    assert meta == expected_metadata(orc.hierarchy, cubic[3] - cubic[0], spec.scale, 32, flags)
