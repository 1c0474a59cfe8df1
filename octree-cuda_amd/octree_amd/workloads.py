"""Synthetic workloads (the named cloud shapes) and the caller-side bounding-box step.

The reference's example clouds (README.md:67-74: morrowbay 119,701,547 x 27 B, heidentor
25,836,417 x 15 B, coin 5,138,448 x 15 B) are external downloads, so every workload here is a
seeded synthetic cloud of the same size and record layout produced by ``libocsynth.so``
(``csrc/synth/cloudgen.cpp``). The bounding box and scale are derived exactly as the reference CLI
does (src/boundingbox.h:10-52, src/argparser.h:149-150: the scale is the float given on the
command line, widened to double).
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from ._lib import load_library


class _OcsParams(ctypes.Structure):
    _fields_ = [
        ("points", ctypes.c_uint64),
        ("cloud_type", ctypes.c_uint32),
        ("box", ctypes.c_double * 3),
        ("clusters", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("offset", ctypes.c_double),
        ("heightfield_fraction", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("threads", ctypes.c_uint32),
    ]


_synth = None


def _lib():
    global _synth
    if _synth is None:
        lib = load_library("libocsynth.so")
        lib.ocs_generate_range.argtypes = [ctypes.POINTER(_OcsParams), ctypes.c_uint64, ctypes.c_uint64,
                                           ctypes.c_void_p]
        lib.ocs_generate_range.restype = ctypes.c_int
        lib.ocs_bounding_box.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
        lib.ocs_bounding_box.restype = ctypes.c_int
        _synth = lib
    return _synth


@dataclasses.dataclass
class CloudSpec:
    """Shape of a synthetic cloud."""

    points: int
    cloud_type: int  # 0 = 3 x float32 + RGB (15 B), 1 = 3 x float64 + RGB (27 B)
    box: tuple = (100.0, 100.0, 20.0)
    clusters: int = 8
    seed: int = 1
    offset: float = 0.0
    heightfield_fraction: float = 0.7
    sigma: float = 2.0

    @property
    def stride(self) -> int:
        return 15 if self.cloud_type == 0 else 27

    @property
    def scale(self) -> float:
        # the CLI passes -d <stride>,<scale> as float (README.md:67-74)
        return float(np.float32(0.001 if self.cloud_type == 0 else 0.01))


@dataclasses.dataclass
class BuildSpec:
    """Processing configuration (ocpi_configure_chunking / ocpi_configure_subsampling)."""

    chunking_grid: int = 512
    merging_threshold: int = 10000
    output_factor: float = 2.2
    subsampling_grid: int = 128
    intra: bool = False
    inter: bool = False
    replacement: bool = True
    random: bool = False
    weighted: bool = False
    seed: int = 0


@dataclasses.dataclass
class Workload:
    name: str
    cloud: CloudSpec
    build: BuildSpec


# The three single-cloud configurations (sizes of the reference's example clouds).
C1 = Workload("coin-like 5,138,448 x 15 B, implicit + inter",
              CloudSpec(5_138_448, 0, (100.0, 100.0, 20.0), 8, 1),
              BuildSpec(512, 10000, 3.0, 128, intra=False, inter=True, random=False))
C2 = Workload("heidentor-like 25,836,417 x 15 B, explicit + intra",
              CloudSpec(25_836_417, 0, (50.0, 50.0, 30.0), 16, 2),
              BuildSpec(512, 10000, 3.0, 128, intra=True, inter=False, random=True))
C3 = Workload("morrowbay-like 119,701,547 x 27 B f64, implicit + inter + weighted",
              CloudSpec(119_701_547, 1, (2000.0, 2000.0, 150.0), 64, 3, offset=1e6),
              BuildSpec(512, 10000, 2.2, 128, intra=False, inter=True, random=False, weighted=True))

WORKLOADS = {"C1": C1, "C2": C2, "C3": C3}


def generate(spec: CloudSpec, first: int = 0, count: int | None = None, threads: int = 0) -> np.ndarray:
    """Returns points [first, first + count) of the cloud as a uint8 array of count * stride bytes."""
    if count is None:
        count = spec.points - first
    out = np.empty(count * spec.stride, dtype=np.uint8)
    p = _OcsParams(spec.points, spec.cloud_type, (ctypes.c_double * 3)(*spec.box), spec.clusters, spec.seed,
                   spec.offset, spec.heightfield_fraction, spec.sigma, threads)
    rc = _lib().ocs_generate_range(ctypes.byref(p), first, count, out.ctypes.data)
    if rc != 0:
        raise ValueError("invalid cloud spec")
    return out


def bounding_box(cloud: np.ndarray, points: int, stride: int, cloud_type: int, threads: int = 0):
    """(real_bb[6], cubic_bb[6]) exactly as src/boundingbox.h computes them."""
    real = (ctypes.c_double * 6)()
    cubic = (ctypes.c_double * 6)()
    rc = _lib().ocs_bounding_box(cloud.ctypes.data, points, stride, cloud_type, threads, real, cubic)
    if rc != 0:
        raise ValueError("bounding box failed")
    return list(real), list(cubic)


def cuboid_cloud(side: int = 128) -> tuple[np.ndarray, list, list]:
    """The reference's cuboid test cloud (tools.cu:7-23, :166-190) as 15-byte records.

    Point i sits at (x + 0.5, y + 0.5, z + 0.5) with x fastest; the reference writes 12-byte
    xyz records, this fixture appends a known colour (r, g, b) = (x, y, z) mod 256.
    Returns (records, cubic_bb, scale) with the reference generator's BB: min 0.5, max side - 0.5.
    """
    n = side ** 3
    idx = np.arange(n, dtype=np.int64)
    z = idx // (side * side)
    y = (idx - z * side * side) // side
    x = (idx - z * side * side) % side
    rec = np.zeros((n, 15), dtype=np.uint8)
    xyz = np.stack([x + 0.5, y + 0.5, z + 0.5], axis=1).astype(np.float32)
    rec[:, :12] = xyz.view(np.uint8).reshape(n, 12)
    rec[:, 12] = (x % 256).astype(np.uint8)
    rec[:, 13] = (y % 256).astype(np.uint8)
    rec[:, 14] = (z % 256).astype(np.uint8)
    bb_max = float(np.float32(side) - np.float32(0.5))
    return rec.reshape(-1), [0.5, 0.5, 0.5, bb_max, bb_max, bb_max], [1.0, 1.0, 1.0]
