"""GPU: the allocation-free fused gather (pm_grid_gather_radiance_noalloc)
returns what pm_grid_gather_radiance returns, bit for bit, over consecutive
calls that share its fixed device scratch: a cell-ordered k = 64 gather at the
1,048,576-query limit, then a smaller large-k gather (three-walk kernel with
its overflow list), then a gather in caller order. Sizes above the limit are
rejected with PM_ERR_UNSUPPORTED before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(t):
    return t.cpu().numpy().view(np.uint32)


def test_noalloc_gather_equals_allocating_gather(ctx):
    import torch

    import pm355
    from pm355.workloads import gather_microbench

    ph, q, qn, aabb = gather_microbench(seed=21, n_photons=2_000_000, n_queries=1 << 20, n_clusters=16)
    g = ctx.grid_build(ph, aabb, cell_size=0.005)
    dq, dn = torch.from_numpy(q).cuda(), torch.from_numpy(qn).cuda()
    S = pm355.KNN_SORT_QUERIES
    # (queries, k, radius, flags), each call smaller than the one before
    for nq, k, r, flags in [(1 << 20, 64, 0.01, S), (50_000, 600, 0.03, S), (30_000, 64, 0.01, 0)]:
        print(f"noalloc gather: nq={nq} k={k} r={r} flags={flags}", flush=True)
        qq, nn = dq[:, :nq].contiguous(), dn[:, :nq].contiguous()
        rgb = torch.full((3, nq), float("nan"), device="cuda")
        cnt = torch.full((nq,), -1, dtype=torch.int32, device="cuda")
        ctx.gather_radiance_noalloc(g, qq, nn, r, k, flags=flags, rgb=rgb, count=cnt)
        ref_rgb, ref_cnt = ctx.gather_radiance(g, qq, nn, r, k, flags=flags, want_count=True)
        ctx.sync()
        assert np.array_equal(_bits(cnt), _bits(ref_cnt))
        assert np.array_equal(_bits(rgb), _bits(ref_rgb))
        assert int(cnt.max()) == k  # the selection narrowed somewhere
    g.close()


def test_noalloc_gather_rejects_sizes_above_the_limit(ctx):
    import torch

    import pm355

    ph = np.zeros((9, 16), np.float32)
    g = ctx.grid_build(ph, np.array([0, 0, 0, 1, 1, 1], np.float32), cell_size=0.5)
    nq = (1 << 20) + 1
    qq = torch.zeros((3, nq), device="cuda")
    rgb = torch.empty((3, nq), device="cuda")
    with pytest.raises(pm355.PMError) as e:
        ctx.gather_radiance_noalloc(g, qq, qq, 0.1, 8, flags=pm355.KNN_SORT_QUERIES, rgb=rgb)
    assert e.value.code == pm355.PM_ERR_UNSUPPORTED
    assert "1048576" in str(e.value)
    g.close()
