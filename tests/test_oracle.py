"""CPU tests that pin the oracle (no GPU needed).

Each check compares the oracle with something it was not written from: the
reference's own fixtures and known-answer values, closed forms, numpy's
independent routines (bincount, stable argsort, cumsum) or brute force.
"""
import os

import numpy as np
import pytest

import pyoracle as O
from conftest import GOLDEN

# Shipped CornellBox-Sphere.obj vertex bounds (min xyz, max xyz).
SHIPPED_AABB = (-1.02, 0.0, -1.04, 1.0, 1.59, 0.99)

# PhotonMapInsertTest.cu:27-34 / PhotonMapGetClosestTest.cu:21-28 photon set.
KD_FIXTURE = np.array([[-10, 7, 20], [-50, 4, 30], [0, 3, 21], [80, 9, 40], [40, 1, 10], [30, 5, 0]],
                      np.float32).T.copy()


def scan_kat_input(n=1_000_000):
    # DynamicArrayTest.cu:160-165: a[i] = i%7 ? (i%3+1)%2 : i%2; a[0] = 512*44.
    i = np.arange(n, dtype=np.int64)
    a = np.where(i % 7 != 0, (i % 3 + 1) % 2, i % 2).astype(np.uint32)
    a[0] = 512 * 44
    return a


def test_scan_kat_closed_forms():
    a = scan_kat_input()
    inc, tot = O.scan_u32(a, inclusive=True)
    # values fixed by the KAT's input formula
    assert tot == 665_385
    assert inc[-1] == 665_385
    assert inc[1023] == 23_186
    assert inc[45055] == 51_492
    assert np.array_equal(inc, np.cumsum(a, dtype=np.uint64).astype(np.uint32))
    exc, tot2 = O.scan_u32(a, inclusive=False)
    assert tot2 == tot
    assert exc[0] == 0 and np.array_equal(exc[1:], inc[:-1])


def test_scan_edge_sizes():
    for n in [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 45057]:
        a = np.random.default_rng(n).integers(0, 1000, n).astype(np.uint32)
        exc, tot = O.scan_u32(a)
        ref = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]).astype(np.uint32) if n else a
        assert np.array_equal(exc, ref)
        assert tot == int(a.astype(np.uint64).sum())


def test_grid_reference_sizing_known_values():
    # PhotonGrid.cu:41-56 for the shipped scene: 1e6 photons / 1000 per cell
    # and 1e5 / 1000.
    rng = np.random.default_rng(0)
    for n, cell_expect, dims_expect in [(1_000_000, 0.186816, (11, 9, 11)), (100_000, 0.402484, (6, 4, 6))]:
        ph = np.zeros((9, n), np.float32)
        lo = np.array(SHIPPED_AABB[:3], np.float32)
        hi = np.array(SHIPPED_AABB[3:], np.float32)
        ph[0:3] = (lo[:, None] + rng.random((3, n), dtype=np.float32) * (hi - lo)[:, None])
        g = O.OracleGrid(ph, SHIPPED_AABB, cell_size=0.0, desired=1000)
        # independent float32 evaluation of the same closed form
        size = np.abs(hi - lo)
        N = (n + 1000 - 1) // 1000
        vol = np.float32(np.float32(size[0] * size[1]) * size[2])
        cell = np.float32(np.cbrt(np.float32(vol / np.float32(N))))
        assert abs(g.cell - float(cell)) <= 1e-6 * float(cell)
        assert abs(g.cell - cell_expect) < 5e-6
        assert tuple(g.dims) == dims_expect
        assert tuple(int(v) for v in np.ceil(size / np.float32(g.cell))) == dims_expect


def test_grid_density_resize():
    # analyze_density (PhotonGrid.cu:120-138): a cell above 5*desired shrinks
    # the cell size to the desired/5 rule.
    n = 20_000
    ph = np.zeros((9, n), np.float32)
    ph[0:3] = 0.5  # every photon in one cell
    g = O.OracleGrid(ph, (0, 0, 0, 1, 1, 1), cell_size=0.0, desired=1000)
    N = (n + 200 - 1) // 200
    assert abs(g.cell - float(np.cbrt(np.float32(1.0 / N)))) < 1e-6


def _numpy_cell_ids(pos, aabb, cell, dims):
    ids = []
    for a in range(3):
        f = (pos[a] - np.float32(aabb[a])) / np.float32(cell)
        f = f.astype(np.float32)
        i = np.where(~(f < dims[a]), dims[a] - 1, np.where(f < 1.0, 0, np.minimum(f.astype(np.int64), dims[a] - 1)))
        ids.append(i.astype(np.int64))
    return ids[0] + ids[1] * dims[0] + ids[2] * dims[0] * dims[1]


def _check_grid_against_numpy(ph, g):
    cid = _numpy_cell_ids(ph[0:3], g.aabb, g.cell, g.dims)
    counts = np.bincount(cid, minlength=g.n_cells).astype(np.uint32)
    assert np.array_equal(g.counts, counts)
    starts = np.concatenate([[0], np.cumsum(counts, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert np.array_equal(g.starts, starts)
    perm = np.argsort(cid, kind="stable").astype(np.uint32)
    assert np.array_equal(g.perm, perm)
    assert np.array_equal(g.photons, ph[:, perm])


def test_grid_fixture_file():
    from pm355.workloads import read_photon_file, write_photon_file

    path = os.path.join(GOLDEN, "global_map_test.bin")
    ph = read_photon_file(path)
    assert ph.shape == (9, 100) and os.path.getsize(path) == 8 + 36 * 100
    assert np.all(ph[3:6] == np.float32(0.01))
    nrm = np.sqrt((ph[6:9].astype(np.float64) ** 2).sum(0))
    assert np.allclose(nrm, 1.0, atol=1e-5)
    aabb = tuple(ph[0:3].min(1)) + tuple(ph[0:3].max(1))
    for cell in [0.0, 0.25, 0.7]:
        g = O.OracleGrid(ph, aabb, cell_size=cell, desired=10)
        _check_grid_against_numpy(ph, g)
    tmp = os.path.join(GOLDEN, "..", "_roundtrip.bin")
    try:
        write_photon_file(tmp, ph)
        assert open(tmp, "rb").read() == open(path, "rb").read()
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)


def test_grid_synthetic_matches_numpy():
    from pm355.workloads import gather_microbench

    ph, q, qn, aabb = gather_microbench(seed=7, n_photons=60_000, n_queries=10, n_clusters=8)
    for cell in [0.01, 0.05, 0.13]:
        g = O.OracleGrid(ph, aabb, cell_size=cell)
        _check_grid_against_numpy(ph, g)


def _sets_equal(i1, d1, c1, i2, d2, c2):
    assert np.array_equal(c1, c2)
    assert np.array_equal(i1, i2)
    assert np.array_equal(d1.view(np.uint32), d2.view(np.uint32))


def test_grid_knn_equals_brute_force():
    from pm355.workloads import gather_microbench

    ph, q, qn, aabb = gather_microbench(seed=3, n_photons=30_000, n_queries=400, n_clusters=4, sigma=0.02)
    # duplicates and an out-of-box query
    ph[0:3, 100:140] = ph[0:3, 99:100]
    q[:, 0] = ph[0:3, 99]
    q[:, 1] = (1.5, 0.5, 0.5)
    for cell, r, k in [(0.05, 0.05, 50), (0.02, 0.05, 64), (0.1, 0.03, 1), (0.05, 0.08, 256)]:
        g = O.OracleGrid(ph, aabb, cell_size=cell)
        gi, gd, gc = g.knn(q, r, k)
        bi, bd, bc = O.knn_brute(g.photons[0:3], q, r, k)
        _sets_equal(gi, gd, gc, bi, bd, bc)
    assert gc[1] == 0 and np.all(gi[1] == 0xFFFFFFFF)


def test_knn_reference_fixture_four_nearest():
    # PhotonMapGetClosestTest.cu:21-52,64: query (80, 9, 40), k = 4 returns
    # photons {3, 4, 5, 2} with distances {0, 50.636, 64.156, 82.444}.
    q = np.array([[80.0], [9.0], [40.0]], np.float32)
    bi, bd, bc = O.knn_brute(KD_FIXTURE, q, 1e4, 4)
    assert bc[0] == 4
    assert sorted(bi[0].tolist()) == [2, 3, 4, 5]
    assert np.allclose(bd[0], [0.0, 50.636, 64.156, 82.444], atol=1e-3)
    ni, npl = O.kdtree_build(KD_FIXTURE)
    ki, kd, kc = O.kdtree_knn(KD_FIXTURE, ni, npl, q, 1e4, 4)
    assert np.array_equal(ki, bi) and np.array_equal(kd, bd)


def test_kdtree_reference_fixture_split_order():
    # PhotonMapInsertTest.cu:36-57: heap nodes of the balanced tree.
    ni, npl = O.kdtree_build(KD_FIXTURE)
    expect = [(30, 5, 0), (-10, 7, 20), (80, 9, 40), (-50, 4, 30), (0, 3, 21), (40, 1, 10)]
    got = [tuple(int(v) for v in KD_FIXTURE[:, i]) for i in ni[:6]]
    assert got == expect
    assert len(ni) == 7 and ni[6] == 0xFFFFFFFF
    assert npl[0] == 0  # x has the largest extent (130 vs 8 and 40)


def _kd_check_invariants(pos, ni, npl):
    n = pos.shape[1]
    used = ni[ni != 0xFFFFFFFF]
    assert sorted(used.tolist()) == list(range(n))  # every photon exactly once
    cap = len(ni)

    def subtree(i):
        out, st = [], [i]
        while st:
            j = st.pop()
            if j < cap and ni[j] != 0xFFFFFFFF:
                out.append(ni[j])
                st += [2 * j + 1, 2 * j + 2]
        return out

    for i in range(cap):
        if ni[i] == 0xFFFFFFFF:
            continue
        p, key = npl[i], (pos[npl[i], ni[i]], ni[i])
        for j in subtree(2 * i + 1):
            assert (pos[p, j], j) < key
        for j in subtree(2 * i + 2):
            assert (pos[p, j], j) > key
        # child plane is the argmax extent of that subtree's points
        for c in (2 * i + 1, 2 * i + 2):
            if c < cap and ni[c] != 0xFFFFFFFF:
                pts = pos[:, subtree(c)]
                ext = np.abs(pts.max(1) - pts.min(1))
                assert npl[c] == int(np.argmax(ext))


def test_kdtree_invariants_with_ties():
    rng = np.random.default_rng(5)
    for n in [1, 2, 3, 5, 7, 8, 33, 200]:
        pos = rng.integers(0, 4, (3, n)).astype(np.float32)  # heavy ties
        ni, npl = O.kdtree_build(pos)
        _kd_check_invariants(pos, ni, npl)


def test_kdtree_knn_equals_brute_force():
    rng = np.random.default_rng(11)
    pos = rng.random((3, 5000), dtype=np.float32)
    pos[:, 10:30] = pos[:, 9:10]
    q = rng.random((3, 300), dtype=np.float32)
    q[:, 0] = pos[:, 9]
    ni, npl = O.kdtree_build(pos)
    for r, k in [(0.05, 10), (0.2, 64), (2.0, 7)]:
        ki, kd, kc = O.kdtree_knn(pos, ni, npl, q, r, k)
        bi, bd, bc = O.knn_brute(pos, q, r, k)
        _sets_equal(ki, kd, kc, bi, bd, bc)


def test_radiance_closed_form():
    # Raytracer.cuh:110-138 evaluated by hand in float64 for two photons.
    ph = np.zeros((9, 2), np.float32)
    ph[0:3, 0] = (0.1, 0.0, 0.0)
    ph[0:3, 1] = (0.0, 0.3, 0.0)
    ph[3:6, 0] = (1.0, 2.0, 3.0)
    ph[3:6, 1] = (0.5, 0.5, 0.5)
    ph[6:9, 0] = (0.0, -1.0, 0.0)  # cos = 1 > 0
    ph[6:9, 1] = (0.0, -1.0, 0.0)
    nrm = np.array([[0.0], [1.0], [0.0]], np.float32)
    idx = np.array([[0, 1]], np.uint32)
    dist = np.array([[0.1, 0.3]], np.float32)
    cnt = np.array([2], np.uint32)
    o32, o64 = O.radiance(idx, dist, cnt, ph, nrm, 2)
    a, b = 1.818, 1.953
    r = float(np.float32(0.3))

    def w(d):
        return a * (1 - (1 - np.exp(-b * d * d / (2 * r * r))) / (1 - np.exp(-b)))

    d0 = float(np.float32(0.1))
    expect = (np.array([1, 2, 3]) * w(d0) + 0.5 * w(r)) / (np.pi * r * r)
    assert np.allclose(o64[:, 0], expect, rtol=1e-6)
    assert np.allclose(o32[:, 0], expect, rtol=1e-5)
    # a photon arriving from behind the surface contributes nothing
    ph[6:9, 1] = (0.0, 1.0, 0.0)
    o32b, o64b = O.radiance(idx, dist, cnt, ph, nrm, 2)
    assert np.allclose(o64b[:, 0], np.array([1, 2, 3]) * w(d0) / (np.pi * r * r), rtol=1e-6)
    # empty set -> black
    o32c, _ = O.radiance(idx, dist, np.array([0], np.uint32), ph, nrm, 2)
    assert np.all(o32c == 0)


def test_radius_threshold_is_exact():
    for r in [0.01, 0.02, 0.05, 0.1, 0.2, 1.0, 3e-5]:
        t = np.float32(O.radius_threshold(r))
        assert np.sqrt(t) <= np.float32(r)
        assert np.sqrt(np.nextafter(t, np.float32(np.inf))) > np.float32(r)


def test_model_aabb_closed_form():
    rng = np.random.default_rng(2)
    tris = rng.random((10, 9), dtype=np.float32) * 4 - 2
    tb = np.array([0, 3, 3, 10], np.uint32)
    out = O.model_aabb(tris, tb)
    for m in (0, 2):
        v = tris[tb[m]:tb[m + 1]].reshape(-1, 3)
        assert np.array_equal(out[m, :3], v.min(0)) and np.array_equal(out[m, 3:], v.max(0))
