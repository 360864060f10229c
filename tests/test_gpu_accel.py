"""GPU: the per-model triangle BVHs return exactly the brute-force nearest
hit, so photon maps and frames are bit-identical with and without them, and
equal to the oracle's (which tests every triangle of every model it enters).

Scenes follow the batched-workload distribution: the open unit box with its
ceiling light plus 1-8 spheres of 200-20k triangles with diffuse, mirror or
glass materials (pm355.scenes.random_scene)."""
import numpy as np
import pytest

import pyoracle as O
from pm355.scenes import cornell_box, cornell_camera, random_scene

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("seed", [3, 8, 21])
def test_trace_bvh_equals_brute_force(ctx, seed):
    sc = random_scene(seed)
    plain, fast = ctx.upload_scene(sc), ctx.upload_scene(sc, accel=True)
    args = (seed, 300_000, 30_000, 10)
    g0, c0, s0 = ctx.trace_photons(plain, *args, max_paths=5_000_000)
    g1, c1, s1 = ctx.trace_photons(fast, *args, max_paths=5_000_000)
    assert s0 == s1
    assert np.array_equal(_np(g0).view(np.uint32), _np(g1).view(np.uint32))
    assert np.array_equal(_np(c0)[:, : s0[2]].view(np.uint32), _np(c1)[:, : s1[2]].view(np.uint32))


@pytest.mark.parametrize("seed,ng,nc", [(5, 4000, 400), (13, 4000, 400), (13, 1000, 20000)])
def test_trace_bvh_matches_oracle(ctx, seed, ng, nc):
    # (1000, 20000): the caustic map fills only in the caustic phase
    sc = random_scene(seed, n_spheres=3, tri_range=(200, 4000))
    fast = ctx.upload_scene(sc, accel=True)
    g, c, st = ctx.trace_photons(fast, seed, ng, nc, 10, max_paths=400_000)
    og, oc, ost = O.trace_photons(sc, seed, ng, nc, 10, max_paths=400_000)
    assert st[1:] == ost[1:]  # paths: the device traces whole batches of paths
    assert np.array_equal(_np(g).view(np.uint32), og.view(np.uint32))
    assert np.array_equal(_np(c)[:, : st[2]].view(np.uint32), oc[:, : st[2]].view(np.uint32))


@pytest.mark.parametrize("seed", [3, 8])
def test_render_bvh_equals_brute_force(ctx, seed):
    sc = random_scene(seed)
    plain, fast = ctx.upload_scene(sc), ctx.upload_scene(sc, accel=True)
    g, c, _ = ctx.trace_photons(fast, seed, 200_000, 20_000, 8, max_paths=5_000_000)
    aabb = sc.aabb()
    gg = ctx.grid_build(g, aabb, cell_size=0.03)
    gc = ctx.grid_build(c, aabb, cell_size=0.03)
    cam = cornell_camera(320, 240)
    r0, u0, h0 = ctx.render_frame(plain, gg, gc, cam, 5, 50, 0.03, 50, 0.03)
    r1, u1, h1 = ctx.render_frame(fast, gg, gc, cam, 5, 50, 0.03, 50, 0.03)
    assert h0 == h1
    assert np.array_equal(_np(r0).view(np.uint32), _np(r1).view(np.uint32))
    assert np.array_equal(_np(u0), _np(u1))


def test_accel_of_another_scene_is_rejected(ctx):
    import pm355

    a = ctx.upload_scene(random_scene(1), accel=True)
    b = ctx.upload_scene(cornell_box())
    b.rec.accel = a.accel.value
    with pytest.raises(pm355.PMError):
        ctx.trace_photons(b, 1, 1000, 100, 4, max_paths=100_000)
    b.rec.accel = None


def _stacked_scene(n_copies=32768):
    """The Cornell box plus a model of n identical triangles (a horizontal
    slab at y = 0.5 under the light). All its BVH boxes coincide, so every
    visited node has both children hit and a traversal pushes at each of the
    tree's ~log2(n) levels: deeper than the 12 stack levels held in LDS, so
    the per-lane overflow entries are used. Identical triangles tie on t, so
    the lowest index must win, as in the brute-force loop."""
    from pm355.scenes import PM_MODEL_DTYPE, Scene, _model

    base = cornell_box()
    tri = np.array([0.2, 0.5, 0.2, 0.8, 0.5, 0.2, 0.5, 0.5, 0.8], np.float32)
    tris = np.tile(tri, (n_copies, 1))
    nrms = np.tile(np.array([0, 1, 0] * 3, np.float32), (n_copies, 1))
    t0 = base.tris9.shape[0]
    models = list(base.models) + [_model(t0, t0 + n_copies, kd=(0.5, 0.5, 0.5), mid=len(base.models))]
    return Scene(np.concatenate([base.tris9, tris]), np.concatenate([base.normals9, nrms]),
                 np.array(models, PM_MODEL_DTYPE), light_model=base.light_model, names=base.names + ["stack"])


def test_deep_bvh_stack_equals_brute_force(ctx):
    sc = _stacked_scene()
    plain, fast = ctx.upload_scene(sc), ctx.upload_scene(sc, accel=True)
    args = (17, 20_000, 2_000, 6)
    g0, c0, s0 = ctx.trace_photons(plain, *args, max_paths=400_000)
    g1, c1, s1 = ctx.trace_photons(fast, *args, max_paths=400_000)
    assert s0 == s1
    assert np.array_equal(_np(g0).view(np.uint32), _np(g1).view(np.uint32))
    assert np.array_equal(_np(c0)[:, : s0[2]].view(np.uint32), _np(c1)[:, : s1[2]].view(np.uint32))
    # photons landed on the slab (its y) from both paths alike
    assert np.any(np.abs(_np(g1)[1] - 0.5) < 1e-6)
    aabb = sc.aabb()
    gg = ctx.grid_build(g1, aabb, cell_size=0.05)
    gc = ctx.grid_build(c1, aabb, cell_size=0.05)
    cam = cornell_camera(96, 64)
    r0, u0, h0 = ctx.render_frame(plain, gg, gc, cam, 3, 20, 0.05, 20, 0.05)
    r1, u1, h1 = ctx.render_frame(fast, gg, gc, cam, 3, 20, 0.05, 20, 0.05)
    assert h0 == h1
    assert np.array_equal(_np(r0).view(np.uint32), _np(r1).view(np.uint32))
