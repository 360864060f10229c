"""GPU: the per-frame pipeline (pm_pipeline_*), which enqueues a whole frame
(trace, both grids rebuilt in place, eye rays, both gathers, fold) with
every count kept on the device.

Its frames equal the eager calls' (pm_trace_photons, pm_grid_build of both
maps, pm_render_frame) for the same seed bit for bit, frame after frame; and
a frame captured into a HIP graph replays with the eager result, several
times, with eager work on another context in between.
"""
import numpy as np
import pytest

from pm355.scenes import cornell_box, cornell_camera

pytestmark = pytest.mark.gpu

NG, NC, CELL, R, K = 200_000, 50_000, 0.02, 0.02, 64


def _eager(ctx, ds, scene, cam, seed, k=K, r=R):
    g, c, st = ctx.trace_photons(ds, seed, NG, NC, 10, max_paths=200_000_000)
    assert st[1:] == (NG, NC)
    gg = ctx.grid_build(g, scene.aabb(), cell_size=CELL)
    gc = ctx.grid_build(c, scene.aabb(), cell_size=CELL)
    rgb, rgb8, hits = ctx.render_frame(ds, gg, gc, cam, 4, k, r, k, r)
    ctx.sync()
    gg.close()
    gc.close()
    return rgb, rgb8, hits


@pytest.mark.parametrize("accel", [True, False], ids=["bvh", "brute"])
def test_pipeline_frames_equal_eager_frames(accel):
    import torch

    import pm355

    scene = cornell_box()
    cam = cornell_camera(160, 120)
    npix = 160 * 120
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx = pm355.Context(0)
        ref_ctx = pm355.Context(0)
        ds = ctx.upload_scene(scene, accel=accel)
        pl = ctx.pipeline(ds, scene.aabb(), CELL, CELL, NG, NC, cam, max_bounces=10, max_paths=200_000_000,
                          plan_seed=5, max_depth=4, k_global=K, r_global=R, k_caustic=K, r_caustic=R)
        for seed in (11, 12, 5):
            print(f"pipeline frame seed={seed} accel={accel}", flush=True)
            rgb8 = torch.zeros(npix * 3, dtype=torch.uint8, device="cuda")
            rgb3 = torch.zeros((3, npix), dtype=torch.float32, device="cuda")
            pl.frame(seed, rgb8, rgb3)
            info = pl.status()
            assert info["complete"] and info["n_global"] == NG and info["n_caustic"] == NC
            e3, e8, hits = _eager(ref_ctx, ds, scene, cam, seed)
            assert info["hits"] == hits
            assert torch.equal(rgb3.view(torch.int32), e3.view(torch.int32)), seed
            assert torch.equal(rgb8, e8), seed
        pl.close()
        ctx.close()
        ref_ctx.close()


def test_pipeline_frame_replays_from_graph():
    import torch

    import pm355

    scene = cornell_box()
    cam = cornell_camera(128, 96)
    npix = 128 * 96
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx = pm355.Context(0)
        ref_ctx = pm355.Context(0)
        ds = ctx.upload_scene(scene, accel=True)
        pl = ctx.pipeline(ds, scene.aabb(), CELL, CELL, NG, NC, cam, max_bounces=10, max_paths=200_000_000,
                          plan_seed=3, max_depth=4, k_global=K, r_global=R, k_caustic=K, r_caustic=R,
                          gather_flags=pm355.KNN_INNER_BOX)
        refs = {}
        for seed in (21, 22):
            e3, e8, hits = _eager(ref_ctx, ds, scene, cam, seed)
            refs[seed] = (e3.clone(), e8.clone())
        s.synchronize()
        outs, graphs = {}, {}
        for seed in (21, 22):
            rgb8 = torch.zeros(npix * 3, dtype=torch.uint8, device="cuda")
            rgb3 = torch.zeros((3, npix), dtype=torch.float32, device="cuda")
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=s):
                rc = ctx.lib.pm_pipeline_frame(pl.h, seed, pm355._ptr(rgb3), pm355._ptr(rgb8))
            assert rc == pm355.PM_OK
            outs[seed], graphs[seed] = (rgb3, rgb8), graph
        for rep in range(3):
            for seed in (22, 21):
                print(f"graph replay rep={rep} seed={seed}", flush=True)
                rgb3, rgb8 = outs[seed]
                rgb3.fill_(-1.0)
                rgb8.zero_()
                _eager(ref_ctx, ds, scene, cam, 99 + rep)  # eager work in between
                graphs[seed].replay()
                s.synchronize()
                assert pl.status()["complete"]
                assert torch.equal(rgb3.view(torch.int32), refs[seed][0].view(torch.int32)), (rep, seed)
                assert torch.equal(rgb8, refs[seed][1]), (rep, seed)
        del graphs
        pl.close()
        ctx.close()
        ref_ctx.close()
