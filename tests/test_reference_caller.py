"""The reference's own caller builds against this library unchanged.

src/main.cpp (the reference's only caller of the session API, src/main.cpp:55-109) is compiled
from where it lies, against include/octreeApi.h and the in-tree libOctreeApi.so instead of the
reference's header and library. Two runs that stop before any device call show that it links
and loads: -h prints the usage, and an invalid coordinate type fails the reference's validation
(src/argparser.h:51-97). The binary is written to pytest's temporary directory, outside the tree.
Runs only where the reference sources are present (this container).
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
MAIN = os.path.join(REFERENCE, "src", "main.cpp")
LIBDIR = os.path.join(REPO, "octree-cuda_amd", "lib")

pytestmark = pytest.mark.skipif(not os.path.exists(MAIN) or shutil.which("g++") is None,
                                reason="the reference sources or g++ are not present")


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libOctreeApi.so")):
        pytest.skip("libOctreeApi.so is not built")
    exe = str(tmp_path_factory.mktemp("caller") / "PotreeConverterGPU")
    cmd = ["g++", "-std=c++17", "-O0", "-w", MAIN,
           "-I", os.path.join(REPO, "include"),
           "-I", os.path.join(REFERENCE, "src"),
           "-I", os.path.join(REFERENCE, "External"),
           "-I", os.path.join(REFERENCE, "External", "spdlog", "include"),
           "-L", LIBDIR, "-lOctreeApi", "-Wl,-rpath," + LIBDIR,
           "-Wl,-rpath-link,/opt/rocm/lib", "-pthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def test_reference_main_prints_usage(caller):
    r = subprocess.run([caller, "-h"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Generates Potree compatible LOD structures on the GPU" in r.stdout
    for flag in ("--file", "--output", "--points", "--type", "--data", "--grids", "--merge_threshold"):
        assert flag in r.stdout


def test_reference_main_rejects_an_invalid_type(caller, tmp_path):
    r = subprocess.run([caller, "-f", str(tmp_path / "missing.bin"), "-o", str(tmp_path / "out"), "-p", "10",
                        "-t", "int", "-d", "15,0.001"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 255, (r.returncode, r.stdout, r.stderr)
    assert 'Data type must be "float" or "double"' in r.stdout + r.stderr
