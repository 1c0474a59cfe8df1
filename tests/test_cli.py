"""The command-line caller bin/octree-amd: options and their validation as the reference CLI
(src/argparser.h:24-112). These cases stop before any device is touched."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "octree-cuda_amd", "bin", "octree-amd")

BASE = ["-f", "cloud.bin", "-o", "out", "-p", "100", "-t", "float", "-d", "15,0.001", "-g", "64,32"]


def run(args):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)


def test_help():
    r = run(["-h"])
    assert r.returncode == 0 and "usage: octree-amd" in r.stdout


@pytest.mark.parametrize("drop,message", [("-f", "No input file"), ("-p", "No point amount"),
                                          ("-t", "No coordinate datatype"), ("-d", "No data info"),
                                          ("-g", "Grid sizes")])
def test_missing_parameters(drop, message):
    i = BASE.index(drop)
    r = run(BASE[:i] + BASE[i + 2:])
    assert r.returncode != 0 and message in r.stderr


@pytest.mark.parametrize("extra,message", [(["-a", "-i"], "Either intra-cell"), (["-w"], "Weighting function"),
                                           (["-t", "half"], "float"), (["-d", "15"], "stride and scale"),
                                           (["-g", "64"], "chunking_grid")])
def test_invalid_parameters(extra, message):
    r = run(BASE + extra)
    assert r.returncode != 0 and message in r.stderr


def test_unreadable_file_fails_cleanly(tmp_path):
    r = run(["-f", str(tmp_path / "missing.bin"), "-o", str(tmp_path / "out")] + BASE[4:])
    assert r.returncode != 0 and "cannot open" in r.stderr
