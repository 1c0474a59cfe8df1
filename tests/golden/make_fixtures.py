"""Writes the golden fixtures of this directory: small seeded clouds, their configuration and the
oracle's complete result (nodes, used size, hierarchy.bin, point LUT and output records).

The reference's own tests hold no data files (only the in-code known answers ported in
tests/test_oracle_kat.py), so these vectors come from the oracle, after it has been pinned by
those known answers, the closed forms and the brute-force checks. They freeze the canonical
results so any later change to the oracle or the GPU path that alters them fails a test.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "octree-cuda_amd"))

import oracle_binding as O  # noqa: E402
from octree_amd import workloads as W  # noqa: E402

# name: (points, cloud type, box, clusters, seed, offset, chunking grid, threshold, factor, subsampling grid,
#        intra, inter, weighted, random, rng seed)
CASES = {
    "f32_inter_implicit": (6000, 0, (12.0, 9.0, 3.0), 3, 21, 0.0, 64, 400, 3.0, 16, False, True, False, False, 0),
    "f32_intra_explicit": (5000, 0, (7.0, 7.0, 2.0), 2, 22, 0.0, 32, 300, 3.0, 16, True, False, False, True, 99),
    "f64_weighted_implicit": (4000, 1, (30.0, 20.0, 5.0), 3, 23, 5e4, 32, 250, 3.0, 8, False, True, True, False, 0),
    "f64_plain_explicit": (7000, 1, (5.0, 5.0, 1.0), 4, 24, 1e6, 64, 500, 3.0, 32, False, False, False, True, 7),
}


def case_spec(name):
    n, t, box, k, seed, off, G, T, F, g, intra, inter, weighted, rnd, rs = CASES[name]
    cloud_spec = W.CloudSpec(n, t, box, k, seed, offset=off)
    build = W.BuildSpec(G, T, F, g, intra=intra, inter=inter, weighted=weighted, random=rnd, seed=rs)
    return cloud_spec, build


def main():
    for name in CASES:
        spec, build = case_spec(name)
        cloud = W.generate(spec, threads=1)
        _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type, threads=1)
        r = O.build(cloud, spec.points, spec.stride, spec.cloud_type, cubic, [spec.scale] * 3, build, threads=1)
        assert r.status == 0, (name, r.status)
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"), cloud=cloud, cubic=np.array(cubic), key=r.key, count=r.count,
            point_count=r.point_count, data_idx=r.data_idx, parent=r.parent, children=r.children, flags=r.flags,
            used=np.array(r.used), lut=r.lut[: r.used], out=r.out[: r.used],
            hierarchy=np.frombuffer(r.hierarchy, np.uint8))
        print(name, spec.points, "points,", len(r.key), "nodes,", r.used, "records")


if __name__ == "__main__":
    main()
