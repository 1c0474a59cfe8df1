"""Locates the in-tree native libraries (built by ``make -C octree-cuda_amd``)."""
from __future__ import annotations

import ctypes
import os

PACKAGE_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PACKAGE_ROOT, "lib")


def library_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def load_library(name: str) -> ctypes.CDLL:
    path = library_path(name)
    if not os.path.exists(path):
        raise RuntimeError(f"{name} is not built; run `make -C {PACKAGE_ROOT}` (or __graft_entry__.build())")
    return ctypes.CDLL(path)
