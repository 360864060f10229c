"""GPU: the batched gather (one persistent launch over many jobs) returns for
every job exactly what the single-job calls return, and what the oracle
returns: jobs of different photon counts, grids (linear and Morton, cell
sizes), radii, k (1 to 2000) and query counts, with and without the
(job, cell) order."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _jobs(seed):
    from pm355.workloads import gather_microbench

    rng = np.random.default_rng(seed)
    specs = []
    for j, (n, nq, r, k, cell_mul, layout) in enumerate([
            (20_000, 3000, 0.05, 50, 1.0, 0), (200_000, 5000, 0.02, 64, 1.0, 0), (5_000, 0, 0.05, 8, 1.0, 0),
            (60_000, 777, 0.03, 1, 0.5, 1), (150_000, 4096, 0.01, 256, 2.0, 0), (10_000, 33, 0.2, 100, 0.5, 0),
            (400_000, 6000, 0.015, 64, 1.0, 1), (1_000, 100, 0.5, 16, 1.0, 0),
            # k above 256 (the reference frame's 400 / 2000): own launches in the same call
            (100_000, 500, 0.05, 400, 1.0, 0), (300_000, 300, 0.08, 2000, 1.0, 1)]):
        ph, q, qn, aabb = gather_microbench(seed=seed * 100 + j, n_photons=n, n_queries=max(nq, 1),
                                            n_clusters=int(rng.integers(1, 16)))
        specs.append((ph, q[:, :nq].copy(), qn[:, :nq].copy(), aabb, r, k, r * cell_mul, layout))
    return specs


@pytest.mark.parametrize("flags", [0, 1, 1 | (1 << 6), 1 << 6], ids=["grid", "grid-sorted", "queue-sorted", "queue"])
def test_batch_equals_single_jobs(ctx, flags):
    specs = _jobs(3)
    grids = [ctx.grid_build(ph, aabb, cell_size=cell, layout=layout) for ph, _, _, aabb, _, _, cell, layout in specs]
    jobs = [(g, q, qn, r, k) for g, (_, q, qn, _, r, k, _, _) in zip(grids, specs)]
    rgb_outs = ctx.gather_batch(jobs, flags=flags)
    knn_outs = ctx.gather_batch([(g, q, None, r, k) for g, q, _, r, k in jobs], flags=flags, want_knn=True)
    ctx.sync()
    for g, (ph, q, qn, aabb, r, k, cell, layout), ro, ko in zip(grids, specs, rgb_outs, knn_outs):
        if q.shape[1] == 0:
            continue
        rgb1, cnt1 = ctx.gather_radiance(g, q, qn, r, k, want_count=True)
        i1, d1, c1 = ctx.knn(g, q, r, k)
        ctx.sync()
        assert np.array_equal(ro["rgb"].cpu().numpy().view(np.uint32), rgb1.cpu().numpy().view(np.uint32))
        assert np.array_equal(ro["count"].cpu().numpy(), cnt1.cpu().numpy())
        assert np.array_equal(ko["idx"].cpu().numpy(), i1.cpu().numpy())
        assert np.array_equal(ko["dist"].cpu().numpy().view(np.uint32), d1.cpu().numpy().view(np.uint32))
        assert np.array_equal(ko["count"].cpu().numpy(), c1.cpu().numpy())


def test_batch_matches_oracle(ctx):
    specs = _jobs(5)
    specs = specs[:4] + specs[8:9]
    grids = [ctx.grid_build(ph, aabb, cell_size=cell, layout=layout) for ph, _, _, aabb, _, _, cell, layout in specs]
    outs = ctx.gather_batch([(g, q, None, r, k) for g, (_, q, _, _, r, k, _, _) in zip(grids, specs)],
                            flags=1, want_knn=True)
    ctx.sync()
    for (ph, q, qn, aabb, r, k, cell, layout), o in zip(specs, outs):
        if q.shape[1] == 0:
            continue
        og = O.OracleGrid(ph, aabb, cell_size=cell, layout=layout)
        oi, od, oc = og.knn(q, r, k)
        assert np.array_equal(o["count"].cpu().numpy().view(np.uint32), oc)
        assert np.array_equal(o["idx"].cpu().numpy().view(np.uint32), oi)
        assert np.array_equal(o["dist"].cpu().numpy().view(np.uint32), od.view(np.uint32))
