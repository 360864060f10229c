"""GPU: calls captured into a HIP graph replay correctly.

The single-pass scans (decoupled look-back) tag their tile words with an
epoch kept in device memory, read by every tile and advanced by each launch,
so a launch replayed from a graph sees a fresh epoch instead of the one
current at capture time. The test captures a cell-ordered fused gather (query
sort: radix passes and scans; then the gather kernel) and an exclusive scan,
then replays the graph several times with eager calls on the same context in
between, checking every replay against the eager results.

Capture rules the test follows: the context is warmed up with the same sizes
first (scratch and look-back state only grow outside a capture), outputs are
allocated before the capture, and the graph holds no memset node (every
clear in these paths is a fill kernel).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gather_and_scan_replay_from_graph():
    import torch

    import pm355
    from pm355 import _ptr
    from pm355.workloads import gather_microbench

    ph, q, qn, aabb = gather_microbench(seed=5, n_photons=300_000, n_queries=20_000, n_clusters=8)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx = pm355.Context(0)  # bound to s
        lib = ctx.lib
        g = ctx.grid_build(ph, aabb, cell_size=0.01)
        dq = torch.from_numpy(q).cuda()
        dn = torch.from_numpy(qn).cuda()
        nq = dq.shape[1]
        flags = pm355.KNN_SORT_QUERIES
        ref = ctx.gather_radiance(g, dq, dn, 0.01, 64, flags=flags)
        xs = torch.from_numpy(np.random.default_rng(3).integers(0, 7, 1_000_003).astype(np.uint32)).cuda()
        sref = torch.empty_like(xs)
        tot = torch.zeros(1, dtype=torch.int64, device="cuda")
        ctx.check(lib.pm_scan_u32(ctx.h, _ptr(xs), _ptr(sref), xs.numel(), 0, _ptr(tot)))
        s.synchronize()
        h = xs.cpu().numpy().astype(np.uint64)
        ex = (np.cumsum(h) - h).astype(np.uint32)
        assert np.array_equal(sref.cpu().numpy(), ex)
        assert int(tot.item()) == int(h.sum())

        rgb = torch.zeros_like(ref)
        sout = torch.zeros_like(xs)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            rc1 = lib.pm_grid_gather_radiance(ctx.h, g.h, _ptr(dq), _ptr(dn), nq, C.c_float(0.01), 64, flags,
                                              _ptr(rgb), None)
            rc2 = lib.pm_scan_u32(ctx.h, _ptr(xs), _ptr(sout), xs.numel(), 0, _ptr(tot))
        assert rc1 == pm355.PM_OK and rc2 == pm355.PM_OK

        for rep in range(4):
            rgb.zero_()
            sout.zero_()
            tot.zero_()
            # eager look-back launches on the same context between replays
            other = ctx.gather_radiance(g, dq, dn, 0.01, 64, flags=flags)
            graph.replay()
            s.synchronize()
            assert torch.equal(other, ref), rep
            assert torch.equal(rgb, ref), f"replay {rep}: gather differs from eager"
            assert torch.equal(sout, sref), f"replay {rep}: scan differs from eager"
            assert int(tot.item()) == int(h.sum())
        del graph
        g.close()
        ctx.close()


@pytest.mark.parametrize("k,r", [(400, 0.05), (2000, 0.08)])
def test_large_k_gather_replays_from_graph(k, r):
    """Radiance gathers above k = 256 (the single-walk 1024-key kernel, and
    the three-walk kernel with its overflow list, whose counter a fill kernel
    clears) replay from a captured graph with the eager results; a cluster
    of exact duplicates sends queries through the overflow list."""
    import torch

    import pm355
    from pm355 import _ptr
    from pm355.workloads import gather_microbench

    ph, q, qn, aabb = gather_microbench(seed=k, n_photons=300_000, n_queries=3_000, n_clusters=8)
    ph[0:3, 1000:2500] = ph[0:3, -1:]
    q[:, 0] = ph[0:3, -1]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx = pm355.Context(0)
        lib = ctx.lib
        g = ctx.grid_build(ph, aabb, cell_size=r / 2)
        dq = torch.from_numpy(q).cuda()
        dn = torch.from_numpy(qn).cuda()
        nq = dq.shape[1]
        ref = ctx.gather_radiance(g, dq, dn, r, k, flags=pm355.KNN_SORT_QUERIES)
        s.synchronize()
        rgb = torch.zeros_like(ref)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            rc = lib.pm_grid_gather_radiance(ctx.h, g.h, _ptr(dq), _ptr(dn), nq, C.c_float(r), k,
                                             pm355.KNN_SORT_QUERIES, _ptr(rgb), None)
        assert rc == pm355.PM_OK
        for rep in range(3):
            rgb.fill_(-1.0)
            other = ctx.gather_radiance(g, dq, dn, r, k, flags=pm355.KNN_SORT_QUERIES)
            graph.replay()
            s.synchronize()
            assert torch.equal(other.nan_to_num(7.0), ref.nan_to_num(7.0)), rep
            assert torch.equal(rgb.nan_to_num(7.0), ref.nan_to_num(7.0)), f"replay {rep} differs from eager"
        del graph
        g.close()
        ctx.close()
