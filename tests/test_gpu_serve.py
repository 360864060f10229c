"""GPU: the continuous-batching server renders every job's frames exactly as
the job alone would, and as the one-frame-at-a-time calls do (photon pass
re-seeded per frame, grids rebuilt, pm_render_frame), whatever else is in
flight and whatever the memory budget admits at a time."""
import numpy as np
import pytest

from pm355.scenes import camera, cornell_box, cornell_camera, random_scene

pytestmark = pytest.mark.gpu

SEED_STEP = 0x9E3779B97F4A7C15


def _jobs(ctx):
    specs = []
    for i, sc in enumerate([cornell_box(), random_scene(4), random_scene(9)]):
        ds = ctx.upload_scene(sc, accel=True)
        cam = [cornell_camera(96, 64), camera((0.5, 0.6, 1.6), (0.1, -0.2, -1.0), (0, 1, 0), 50.0, 64, 48),
               cornell_camera(80, 60)][i]
        specs.append(dict(dscene=ds, aabb=sc.aabb(), cell_size=(0.05, 0.04, 0.03)[i],
                          n_global=(50_000, 30_000, 20_000)[i], n_caustic=(5_000, 3_000, 0)[i], camera=cam,
                          frames=(3, 2, 4)[i], seed=100 + i, max_depth=(4, 3, 5)[i], k_global=(50, 64, 32)[i],
                          r_global=(0.05, 0.04, 0.03)[i], k_caustic=(40, 20, 16)[i], r_caustic=(0.05, 0.04, 0.03)[i],
                          max_paths=20_000_000))
    return specs


def _one_at_a_time(ctx, s):
    out = []
    for f in range(s["frames"]):
        seed = (s["seed"] + SEED_STEP * (f + 1)) & ((1 << 64) - 1)
        g, c, st = ctx.trace_photons(s["dscene"], seed, s["n_global"], s["n_caustic"], 10, max_paths=s["max_paths"])
        # the maps' first st[1] / st[2] photons (all of them once the paths fill them)
        gg = ctx.grid_build(g[:, :st[1]], s["aabb"], cell_size=s["cell_size"])
        gc = ctx.grid_build(c[:, :st[2]], s["aabb"], cell_size=s["cell_size"]) if st[2] > 0 else None
        _, rgb8, _ = ctx.render_frame(s["dscene"], gg, gc, s["camera"], s["max_depth"], s["k_global"],
                                      s["r_global"], s["k_caustic"], s["r_caustic"])
        ctx.sync()
        out.append(rgb8.cpu().numpy())
    return np.stack(out)


def _serve(ctx, specs, budget=0):
    srv = ctx.server(budget)
    outs = [srv.submit(**s)[1] for s in specs]
    info = srv.info()
    steps = 0
    while srv.info()["finished"] < len(specs):
        srv.step()
        steps += 1
        assert steps < 100
    ctx.sync()
    lat = srv.latencies()
    res = [o.cpu().numpy() for o in outs], info, lat, srv.info()
    srv.close()
    return res


def test_served_frames_equal_alone_and_one_at_a_time(ctx):
    specs = _jobs(ctx)
    mixed, info0, lat, info = _serve(ctx, specs)
    assert info0["queued"] == 3 and info["finished"] == 3 and info["frames"] == sum(s["frames"] for s in specs)
    assert len(lat) == info["frames"] and np.all(lat > 0)
    for s, m in zip(specs, mixed):
        alone, _, _, _ = _serve(ctx, [s])
        assert np.array_equal(m, alone[0])
        assert np.array_equal(m, _one_at_a_time(ctx, s))
        assert m.any()


def test_budget_admits_jobs_as_memory_frees(ctx):
    specs = _jobs(ctx)
    small, info0, _, _ = _serve(ctx, specs, budget=100 << 20)  # one job at a time
    wide, _, _, _ = _serve(ctx, specs)
    for a, b in zip(small, wide):
        assert np.array_equal(a, b)


def _dark_scene():
    """A light facing away from the only other surface: every photon
    escapes, so a trace stores nothing."""
    from pm355.scenes import PM_MODEL_DTYPE, Scene, _model, _quad

    P = lambda *a: np.array(a, np.float32)
    lt, ln = _quad(P(0.4, 0.5, 0.4), P(0.6, 0.5, 0.4), P(0.6, 0.5, 0.6), P(0.4, 0.5, 0.6), P(0, -1, 0))
    ct, cn = _quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1), P(0, -1, 0))
    models = np.array([_model(0, 2, kd=(0.8, 0.8, 0.8), ke=(1, 1, 1), mid=0), _model(2, 4, kd=(0.7, 0.7, 0.7), mid=1)],
                      PM_MODEL_DTYPE)
    return Scene(np.concatenate([lt, ct]), np.concatenate([ln, cn]), models, light_model=0)


def test_empty_photon_map(ctx):
    """A trace that stores no photon gives an empty grid: gathers find
    nothing and estimate zero, and a served job over it renders the same
    frames as the calls one at a time (emission only)."""
    import torch

    aabb = (0, 0, 0, 1, 1, 1)
    empty = ctx.grid_build(torch.zeros((9, 0), dtype=torch.float32, device="cuda"), aabb, cell_size=0.05)
    assert empty.info.n_photons == 0
    q = torch.rand((3, 500), device="cuda")
    qn = torch.zeros_like(q)
    qn[1] = 1
    idx, dist, cnt = ctx.knn(empty, q, 0.05, 16)
    rgb, c2 = ctx.gather_radiance(empty, q, qn, 0.05, 16, want_count=True)
    ctx.sync()
    assert not cnt.any() and not c2.any() and not rgb.any()
    assert np.all(idx.cpu().numpy().view(np.uint32) == 0xFFFFFFFF)
    sc = _dark_scene()
    ds = ctx.upload_scene(sc)
    cam = camera((0.5, 0.1, 1.2), (0.0, 0.6, -1.0), (0, 1, 0), 60.0, 64, 48)
    spec = dict(dscene=ds, aabb=sc.aabb(), cell_size=0.05, n_global=1000, n_caustic=100, camera=cam, frames=2,
                seed=5, max_depth=3, k_global=16, r_global=0.05, k_caustic=8, r_caustic=0.05, max_paths=20_000)
    got = _serve(ctx, [spec])[0][0]
    for f in range(2):
        seed = (spec["seed"] + SEED_STEP * (f + 1)) & ((1 << 64) - 1)
        g, c, st = ctx.trace_photons(ds, seed, 1000, 100, 10, max_paths=20_000)
        assert st[1] == 0 and st[2] == 0
        gg = ctx.grid_build(g[:, :0], spec["aabb"], cell_size=0.05)
        _, rgb8, _ = ctx.render_frame(ds, gg, None, cam, 3, 16, 0.05, 8, 0.05)
        ctx.sync()
        assert np.array_equal(got[f], rgb8.cpu().numpy())
    assert got.any()  # the light is in view


def _serve_staggered(ctx, specs, arrivals, workers=0):
    """Submit job i at step arrivals[i]; step until every job is done."""
    srv = ctx.server(0)
    if workers:
        srv.set_workers(workers)
    outs = [None] * len(specs)
    steps = 0
    while True:
        for i, a in enumerate(arrivals):
            if a == steps:
                outs[i] = srv.submit(**specs[i])[1]
        if steps > max(arrivals) and srv.info()["finished"] == len(specs):
            break
        srv.step()
        steps += 1
        assert steps < 100
    ctx.sync()
    res = [o.cpu().numpy() for o in outs]
    srv.close()
    return res


def test_maps_short_of_their_budget(ctx):
    """Jobs whose path limit stops the photon pass before the maps fill
    (their blocks keep the plane stride of the requested counts): served
    frames equal the one-at-a-time frames built from the photons stored,
    with and without worker contexts."""
    specs = _jobs(ctx)
    for s in specs:
        s["max_paths"] = s["n_global"] // 4
    for s in specs:
        seed = (s["seed"] + SEED_STEP) & ((1 << 64) - 1)
        _, _, st = ctx.trace_photons(s["dscene"], seed, s["n_global"], s["n_caustic"], 10, max_paths=s["max_paths"])
        assert 0 < st[1] < s["n_global"]
    alone = [_one_at_a_time(ctx, s) for s in specs]
    for workers in (0, 2):
        got = _serve_staggered(ctx, specs, arrivals=[0, 0, 1], workers=workers)
        for a, m in zip(alone, got):
            assert np.array_equal(m, a), workers


def test_staggered_arrivals_equal_alone(ctx):
    """Jobs arriving at different steps render every frame exactly as the
    job alone."""
    specs = _jobs(ctx)
    mixed = _serve_staggered(ctx, specs, arrivals=[0, 1, 3])
    for s, m in zip(specs, mixed):
        print("staggered job", s["seed"], flush=True)
        alone = _serve(ctx, [s])[0][0]
        assert np.array_equal(m, alone)
        assert m.any()


def test_staggered_arrivals_with_workers_equal_alone(ctx):
    """Jobs arriving at different steps, served with three worker contexts
    (their trace / grid / eye-ray stages concurrent on three streams),
    render every frame exactly as the job alone on a server without
    workers."""
    specs = _jobs(ctx)
    mixed = _serve_staggered(ctx, specs, arrivals=[0, 1, 3], workers=3)
    for s, m in zip(specs, mixed):
        print("staggered job", s["seed"], flush=True)
        alone = _serve(ctx, [s])[0][0]
        assert np.array_equal(m, alone)
        assert m.any()


def test_tuned_jobs_equal_alone_and_mixed():
    """Jobs flagged PM_KNN_TUNED build their grids at the tuned cell size and
    gather with the tuned variant; they render the same frames alone and in a
    staggered mix."""
    import pm355
    from pm355.workloads import gather_microbench

    c = pm355.Context(0)
    try:
        ph, q, qn, aabb = gather_microbench(seed=5, n_photons=200_000, n_queries=20_000)
        t = c.tune_gather(ph, aabb, q, qn, 0.05, 50)
        assert t["found"] == 1
        specs = _jobs(c)
        for s in specs:
            s["gather_flags"] = pm355.KNN_TUNED
        mixed = _serve_staggered(c, specs, arrivals=[0, 0, 2])
        for s, m in zip(specs, mixed):
            alone = _serve(c, [s])[0][0]
            assert np.array_equal(m, alone)
            assert m.any()
    finally:
        c.close()
