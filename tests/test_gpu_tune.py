"""GPU: the gather autotuner (pm_tune_gather / pm_tuning_lookup /
PM_KNN_TUNED). Its choice is one of the candidates, no slower than the
default it is measured against, remembered per (k, photon density), and a
gather flagged PM_KNN_TUNED runs exactly the chosen variant: bit-identical to
the same gather with the chosen bits given explicitly. Per-job flags of a
batched gather select the variant of that job alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRACS = (0.25, 1.0 / 3.0, 0.4, 0.5, 2.0 / 3.0, 1.0)


def _data():
    from pm355.workloads import gather_microbench

    # 1M photons at r = 0.04: the photons within r of the 16M / r = 0.01 set
    ph, q, qn, aabb = gather_microbench(seed=21, n_photons=1_000_000, n_queries=60_000)
    return ph, q, qn, aabb, 0.04, 64


def test_tuning_choice_lookup_and_tuned_flag():
    import torch

    import pm355

    ph, q, qn, aabb, r, k = _data()
    ctx = pm355.Context(0)
    try:
        assert ctx.tuning_lookup(ph.shape[1], aabb, r, k)["found"] == 0
        t = ctx.tune_gather(ph, aabb, q, qn, r, k)
        print("tuning", t, flush=True)
        assert t["found"] == 1
        assert min(abs(t["cell_over_r"] - f) for f in FRACS) < 1e-6
        assert t["flags"] in set(pm355.KNN_VARIANTS.values())
        assert 0.0 < t["ms"] <= t["ms_default"]
        # remembered for this k and density (also a nearby density), not for another k
        lk = ctx.tuning_lookup(ph.shape[1], aabb, r, k)
        assert lk["found"] == 1 and lk["flags"] == t["flags"] and lk["cell_over_r"] == t["cell_over_r"]
        assert ctx.tuning_lookup(ph.shape[1] // 2, aabb, r, k)["found"] == 1
        other = ctx.tuning_lookup(ph.shape[1], aabb, r, 32)
        assert other["found"] == 0 and other["cell_over_r"] == 1.0 and other["flags"] == 0
        # PM_KNN_TUNED runs the chosen variant
        d_ph = torch.from_numpy(ph).cuda()
        g = ctx.grid_build(d_ph, aabb, cell_size=r * t["cell_over_r"])
        sort = pm355.KNN_SORT_QUERIES
        rgb_t, cnt_t = ctx.gather_radiance(g, q, qn, r, k, flags=sort | pm355.KNN_TUNED, want_count=True)
        rgb_e, cnt_e = ctx.gather_radiance(g, q, qn, r, k, flags=sort | t["flags"], want_count=True)
        rgb_d, cnt_d = ctx.gather_radiance(g, q, qn, r, k, flags=sort, want_count=True)
        ctx.sync()
        assert torch.equal(rgb_t.view(torch.int32), rgb_e.view(torch.int32))
        assert torch.equal(cnt_t, cnt_e) and torch.equal(cnt_t, cnt_d)
        # other variants select the same neighbours; the radiance sums them in
        # another order
        torch.testing.assert_close(rgb_t, rgb_d, rtol=1e-5, atol=1e-12)
        # without a remembered tuning for this k the flag changes nothing
        rgb_n = ctx.gather_radiance(g, q, qn, r, 32, flags=sort | pm355.KNN_TUNED)
        rgb_n0 = ctx.gather_radiance(g, q, qn, r, 32, flags=sort)
        ctx.sync()
        assert torch.equal(rgb_n.view(torch.int32), rgb_n0.view(torch.int32))
        g.close()
    finally:
        ctx.close()


def test_batched_job_flags_select_the_job_variant():
    import torch

    import pm355

    ph, q, qn, aabb, r, k = _data()
    ctx = pm355.Context(0)
    try:
        d_ph = torch.from_numpy(ph).cuda()
        g = ctx.grid_build(d_ph, aabb, cell_size=r * 0.4)
        qa, na = q[:, :30_000], qn[:, :30_000]
        qb, nb = q[:, 30_000:], qn[:, 30_000:]
        outs = ctx.gather_batch([(g, qa, na, r, k, pm355.KNN_INNER_BOX), (g, qb, nb, r, k, 0)],
                                flags=pm355.KNN_SORT_QUERIES)
        ra = ctx.gather_radiance(g, qa, na, r, k, flags=pm355.KNN_INNER_BOX)
        rb = ctx.gather_radiance(g, qb, nb, r, k, flags=0)
        ctx.sync()
        assert torch.equal(outs[0]["rgb"].view(torch.int32), ra.view(torch.int32))
        assert torch.equal(outs[1]["rgb"].view(torch.int32), rb.view(torch.int32))
        g.close()
    finally:
        ctx.close()
