"""The order-free per-node record comparison used by the full-size GPU tests, checked on the CPU
with oracle builds: records shuffled inside their nodes compare equal; a record moved to another
node, another point index and a changed colour are each caught."""
import numpy as np
import pytest

import oracle_binding as O
from octree_amd import workloads as W
from parity import compare_record_sets, synthetic


@pytest.fixture(scope="module")
def built():
    cloud, spec, cubic = synthetic(40_000, 0, seed=21)
    build = W.BuildSpec(64, 2000, 3.0, 32, intra=False, inter=True, random=False)
    orc = O.build(cloud, spec.points, spec.stride, 0, cubic, [spec.scale] * 3, build)
    assert orc.status == 0
    return orc


def _fake_gpu(orc, lut, out):
    return ({"used": orc.used}, None, lut, out)


def _shuffled(orc, rng):
    used = orc.used
    lut, out = orc.lut[:used].copy(), orc.out[:used].copy()
    for i in np.nonzero(orc.finished)[0]:
        a, n = int(orc.data_idx[i]), int(orc.point_count[i])
        p = a + rng.permutation(n)
        lut[a:a + n], out[a:a + n] = lut[p].copy(), out[p].copy()
    return lut, out


def test_shuffled_records_compare_equal(built):
    orc = built
    lut, out = _shuffled(orc, np.random.default_rng(1))
    fin = orc.finished
    assert compare_record_sets(_fake_gpu(orc, lut, out), orc, fin & ~orc.internal, fin & orc.internal) == fin.sum()


def test_moved_record_is_caught(built):
    orc = built
    lut, out = _shuffled(orc, np.random.default_rng(2))
    leaves = np.nonzero(orc.finished & ~orc.internal & (orc.point_count > 0))[0]
    a, b = int(orc.data_idx[leaves[0]]), int(orc.data_idx[leaves[1]])
    lut[a], lut[b] = lut[b], lut[a]  # two points change nodes, counts stay
    out[a], out[b] = out[b].copy(), out[a].copy()
    with pytest.raises(AssertionError, match="hold other records"):
        compare_record_sets(_fake_gpu(orc, lut, out), orc, orc.finished, np.zeros_like(orc.finished))


def test_wrong_index_and_colour_are_caught(built):
    orc = built
    fin = orc.finished
    internal = np.nonzero(orc.internal)[0]
    pos = int(orc.data_idx[internal[0]])
    lut, out = orc.lut[:orc.used].copy(), orc.out[:orc.used].copy()
    lut[pos] += 1
    with pytest.raises(AssertionError, match="hold other records"):
        compare_record_sets(_fake_gpu(orc, lut, out), orc, fin, np.zeros_like(fin))
    lut[pos] -= 1
    out["r"][pos] ^= 1
    with pytest.raises(AssertionError, match="hold other records"):
        compare_record_sets(_fake_gpu(orc, lut, out), orc, fin, np.zeros_like(fin))
    # a colour change is invisible where colours are left out
    assert compare_record_sets(_fake_gpu(orc, lut, out), orc, fin & ~orc.internal, fin & orc.internal) == fin.sum()
