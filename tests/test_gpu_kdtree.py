"""GPU parity: photon kd-tree split order and model AABBs vs the oracle."""
import numpy as np
import pytest

import pyoracle as O
from test_oracle import KD_FIXTURE

pytestmark = pytest.mark.gpu


def _photons(pos):
    n = pos.shape[1]
    ph = np.zeros((9, n), np.float32)
    ph[0:3] = pos
    ph[3:6] = 1.0
    ph[6:9] = 1.0
    return ph


def _check(ctx, pos):
    t = ctx.kdtree_build(_photons(pos))
    ni, npl = t.to_host()
    oi, opl = O.kdtree_build(pos)
    assert t.desc.n_slots == len(oi)
    assert np.array_equal(ni, oi), np.nonzero(ni != oi)[0][:10]
    assert np.array_equal(npl, opl)
    return t


def test_kdtree_reference_fixture(ctx):
    t = _check(ctx, KD_FIXTURE)
    ni, _ = t.to_host()
    got = [tuple(int(v) for v in KD_FIXTURE[:, i]) for i in ni[:6]]
    assert got == [(30, 5, 0), (-10, 7, 20), (80, 9, 40), (-50, 4, 30), (0, 3, 21), (40, 1, 10)]


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 9, 100, 1023, 1024, 4097])
def test_kdtree_ties(ctx, n):
    pos = np.random.default_rng(n).integers(0, 5, (3, n)).astype(np.float32)
    pos[0, : n // 3] = -0.0  # signed zeros tie with +0
    _check(ctx, pos)


def test_kdtree_synthetic(ctx):
    from pm355.workloads import gather_microbench

    ph, _, _, _ = gather_microbench(seed=21, n_photons=300_000, n_queries=1, n_clusters=8)
    _check(ctx, ph[0:3])


def test_model_aabb(ctx):
    rng = np.random.default_rng(3)
    tris = (rng.random((5000, 9), dtype=np.float32) * 6 - 3)
    tb = np.array([0, 10, 10, 1000, 4999, 5000], np.uint32)
    tb_ok = np.array([0, 10, 1000, 4999, 5000], np.uint32)
    for bounds in (tb, tb_ok):
        got = ctx.model_aabb(tris, bounds).cpu().numpy()
        ref = O.model_aabb(tris, bounds)
        assert np.array_equal(got, ref)
