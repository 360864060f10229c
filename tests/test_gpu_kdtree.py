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


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 9, 100, 1023, 1024, 1025, 2047, 2048, 2049, 4097, 65537])
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


def _kd_knn_check(ctx, pos, q, r, k):
    t = ctx.kdtree_build(_photons(pos))
    gi, gd, gc = ctx.kdtree_knn(t, q, r, k)
    gi = gi.cpu().numpy().view(np.uint32)
    gd = gd.cpu().numpy()
    gc = gc.cpu().numpy().view(np.uint32)
    ni, npl = O.kdtree_build(pos)
    oi, od, oc = O.kdtree_knn(pos, ni, npl, q, r, k)
    assert np.array_equal(gc, oc)
    assert np.array_equal(gi, oi), np.nonzero((gi != oi).any(axis=1))[0][:10]
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))
    return gi, gd, gc


def test_kdtree_knn_reference_fixture(ctx):
    # PhotonMapGetClosestTest.cu:21-52,64: query (80, 9, 40), k = 4 returns
    # photons {3, 4, 5, 2} with distances {0, 50.636, 64.156, 82.444}.
    q = np.array([[80.0], [9.0], [40.0]], np.float32)
    gi, gd, gc = _kd_knn_check(ctx, KD_FIXTURE, q, 1e4, 4)
    assert gc[0] == 4 and sorted(gi[0].tolist()) == [2, 3, 4, 5]
    assert np.allclose(gd[0], [0.0, 50.636, 64.156, 82.444], atol=1e-3)


@pytest.mark.parametrize("n,nq,r,k", [(1, 5, 2.0, 3), (9, 40, 0.5, 4), (5000, 700, 0.05, 16),
                                      (60000, 2000, 0.03, 64), (60000, 500, 0.08, 256), (200000, 3000, 0.01, 100),
                                      (400000, 500, 0.08, 256)])
def test_kdtree_knn_matches_oracle_and_brute(ctx, n, nq, r, k):
    rng = np.random.default_rng(n + k)
    pos = rng.random((3, n), dtype=np.float32)
    if n > 40:
        pos[:, 10:40] = pos[:, 9:10]  # exact duplicates: ties on d2 resolved by index
        pos[1, 100:n // 4] = 0.0  # a plane of photons
    q = rng.random((3, nq), dtype=np.float32)
    q[:, 0] = pos[:, min(9, n - 1)]
    gi, gd, gc = _kd_knn_check(ctx, pos, q, r, k)
    if n * nq <= 2e8:
        bi, bd, bc = O.knn_brute(pos, q, r, k)
        assert np.array_equal(gc, bc) and np.array_equal(gi, bi)


def test_kdtree_knn_ties(ctx):
    pos = np.random.default_rng(4).integers(0, 6, (3, 3000)).astype(np.float32) * 0.01
    q = np.random.default_rng(5).integers(0, 6, (3, 400)).astype(np.float32) * 0.01
    _kd_knn_check(ctx, pos, q, 0.02, 50)


def test_kdtree_knn_deep_stack(ctx):
    """A 21-level tree (2^20 photons) searched with r = 0.1 (about 4,200
    photons within r per query, 520 within the inner half-radius ball): the
    wave's node stack holds hundreds of pending subtrees, and the key buffer
    narrows many times, its bound then pruning subtrees. Results equal the
    oracle's kd search and brute force."""
    rng = np.random.default_rng(77)
    pos = rng.random((3, 1 << 20), dtype=np.float32)
    q = rng.random((3, 300), dtype=np.float32)
    gi, gd, gc = _kd_knn_check(ctx, pos, q, 0.1, 64)
    bi, bd, bc = O.knn_brute(pos, q, 0.1, 64)
    assert np.array_equal(gc, bc) and np.array_equal(gi, bi)
    assert np.all(gc == 64)
