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


def test_batched_pipeline_frames_replay_from_one_graph():
    """Three scenes' frames (three pipelines, one per context, on one stream)
    captured into a single HIP graph: every replay renders each scene's frame
    exactly as the eager calls do, with eager work in between replays."""
    import torch

    import pm355
    from pm355.scenes import camera, random_scene

    scenes = [cornell_box(), random_scene(4), random_scene(9)]
    cams = [cornell_camera(96, 72), camera((0.5, 0.6, 1.6), (0.1, -0.2, -1.0), (0, 1, 0), 50.0, 80, 64),
            cornell_camera(64, 48)]
    seeds = (31, 32, 33)
    ncs = (NC // 4, NC // 4, 0)  # the third frame without a caustic map
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ref_ctx = pm355.Context(0)
        ctxs, pls, dss, outs = [], [], [], []
        for sc, cam, nc in zip(scenes, cams, ncs):
            c = pm355.Context(0)
            ds = c.upload_scene(sc, accel=True)
            pl = c.pipeline(ds, sc.aabb(), CELL, CELL, NG // 4, nc, cam, max_bounces=10,
                            max_paths=200_000_000, plan_seed=2, max_depth=4, k_global=K, r_global=R, k_caustic=K,
                            r_caustic=R)
            npix = int(cam["width"].item()) * int(cam["height"].item())
            outs.append((torch.zeros((3, npix), dtype=torch.float32, device="cuda"),
                         torch.zeros(npix * 3, dtype=torch.uint8, device="cuda")))
            ctxs.append(c)
            pls.append(pl)
            dss.append(ds)
        refs = []
        for ds, sc, cam, seed, nc in zip(dss, scenes, cams, seeds, ncs):
            g, c_, st = ref_ctx.trace_photons(ds, seed, NG // 4, nc, 10, max_paths=200_000_000)
            gg = ref_ctx.grid_build(g[:, :st[1]], sc.aabb(), cell_size=CELL)
            gc = ref_ctx.grid_build(c_[:, :st[2]], sc.aabb(), cell_size=CELL) if st[2] > 0 else None
            e3, e8, _ = ref_ctx.render_frame(ds, gg, gc, cam, 4, K, R, K, R)
            ref_ctx.sync()
            refs.append((e3.clone(), e8.clone()))
        s.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            for c, pl, seed, (o3, o8) in zip(ctxs, pls, seeds, outs):
                assert c.lib.pm_pipeline_frame(pl.h, seed, pm355._ptr(o3), pm355._ptr(o8)) == pm355.PM_OK
        for rep in range(3):
            print(f"batched graph replay rep={rep}", flush=True)
            for o3, o8 in outs:
                o3.fill_(-1.0)
                o8.zero_()
            _eager(ref_ctx, dss[0], scenes[0], cams[0], 90 + rep)  # eager work in between
            graph.replay()
            s.synchronize()
            for i, (pl, (o3, o8), (e3, e8)) in enumerate(zip(pls, outs, refs)):
                assert pl.status()["complete"], i
                assert torch.equal(o3.view(torch.int32), e3.view(torch.int32)), (rep, i)
                assert torch.equal(o8, e8), (rep, i)
        del graph
        for pl, c in zip(pls, ctxs):
            pl.close()
            c.close()
        ref_ctx.close()
