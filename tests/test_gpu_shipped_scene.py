"""GPU parity on the reference's own scene: CornellBox-Sphere (OBJ/MTL from
the reference's Models/cornell_box_sphere, loaded by pm355.objload) with the
reference's camera and frame parameters (kernel.cu:14-30, Initializer.cu:
145-146, Raytracer.cuh:187, 200).

Photon maps are compared bit for bit with the oracle, up to the reference's
full 1M global + 100k caustic photons; frames with the reference's k and
radii (2000 / 0.1 global, 400 / 0.2 caustic) are compared on a reduced
canvas under the same bounds as the synthetic-scene frame tests (summation
order of the radiance sums, device vs host expf / powf).
"""
import os

import numpy as np
import pytest

import pyoracle as O
from conftest import GOLDEN
from pm355.objload import load_obj
from pm355.scenes import REFERENCE_FRAME as RF
from pm355.scenes import reference_camera

pytestmark = pytest.mark.gpu

RMSE_MAX = 1e-5      # HDR-mapped float image, per-pixel RMSE over the frame
ABS_MAX = 1e-4       # any channel, plus the HDR slope term below
U8_STEP_FRAC = 1e-3  # fraction of uchar channels allowed to differ by 1
EXP_ULPS = 4         # expf / powf ulps between the device and host libraries


def _channel_bound(m):
    """Per-channel bound on |device - oracle| of the HDR-mapped value
    m = (1 - e)^(1/2.2), e = expf(-x / 0.6) (Raytracer.cuh:95-107). An ulp
    of e (2^-24 below 1) moves m by 2^-24 / (2.2 m^1.2), which passes 1e-4
    on dim pixels (m below about 1e-3, x below about 1e-6), where 1 - e
    cancels."""
    m = np.maximum(m.astype(np.float64), 1e-6)
    return ABS_MAX + EXP_ULPS * 2.0 ** -24 / (2.2 * m ** 1.2)


@pytest.fixture(scope="module")
def scene():
    return load_obj(os.path.join(GOLDEN, "cornell_box_sphere"), "CornellBox-Sphere.obj")


@pytest.fixture(scope="module", params=[False, True], ids=["brute", "bvh"])
def dscene(ctx, scene, request):
    return ctx.upload_scene(scene, accel=request.param)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("seed,ng,nc", [(1, 20000, 2000), (9, 3000, 6000)])
def test_shipped_trace_parity(ctx, scene, dscene, seed, ng, nc):
    print(f"shipped scene trace seed={seed} ng={ng} nc={nc} accel={dscene.rec.accel is not None}", flush=True)
    g, c, st = ctx.trace_photons(dscene, seed, ng, nc, RF["max_bounces"], max_paths=20_000_000)
    og, oc, ost = O.trace_photons(scene, seed, ng, nc, RF["max_bounces"], max_paths=20_000_000)
    assert st[1:] == ost[1:]
    assert np.array_equal(_np(g).view(np.uint32), og.view(np.uint32))
    assert np.array_equal(_np(c).view(np.uint32), oc.view(np.uint32))


@pytest.fixture(scope="module")
def reference_maps(ctx, scene):
    """The reference's photon maps (1M global, 100k caustic), GPU and oracle."""
    ds = ctx.upload_scene(scene, accel=True)
    g, c, st = ctx.trace_photons(ds, 2, RF["n_global"], RF["n_caustic"], RF["max_bounces"])
    og, oc, ost = O.trace_photons(scene, 2, RF["n_global"], RF["n_caustic"], RF["max_bounces"])
    return ds, g, c, st, og, oc, ost


def test_shipped_reference_photon_maps(reference_maps):
    ds, g, c, st, og, oc, ost = reference_maps
    assert st[1:] == ost[1:] == (RF["n_global"], RF["n_caustic"])
    assert np.array_equal(_np(g).view(np.uint32), og.view(np.uint32))
    assert np.array_equal(_np(c).view(np.uint32), oc.view(np.uint32))


@pytest.mark.parametrize("W,H", [(160, 120), (256, 192)])
def test_shipped_reference_frame(ctx, scene, reference_maps, W, H):
    ds, g, c, st, og, oc, ost = reference_maps
    aabb = scene.aabb()
    gg = ctx.grid_build(g, aabb)  # the reference's sizing from the photon count
    gc = ctx.grid_build(c, aabb)
    cam = reference_camera(W, H)
    print(f"shipped scene frame {W}x{H}", flush=True)
    rgb, rgb8, hits = ctx.render_frame(ds, gg, gc, cam, RF["max_depth"], RF["k_global"], RF["r_global"],
                                       RF["k_caustic"], RF["r_caustic"])
    ctx.sync()
    orgb, orgb8, ohits = O.render(scene, O.OracleGrid(og, aabb), O.OracleGrid(oc, aabb), cam, RF["max_depth"],
                                  RF["k_global"], RF["r_global"], RF["k_caustic"], RF["r_caustic"])
    assert hits == ohits
    got = _np(rgb)
    err = np.abs(got.astype(np.float64) - orgb)
    rmse = np.sqrt(np.mean(err ** 2))
    worst = np.unravel_index(np.argmax(err), err.shape)
    d8 = np.abs(_np(rgb8).astype(np.int32) - orgb8.astype(np.int32))
    diag = (f"rmse {rmse:.3g}, max abs {err.max():.3g} at {worst} (device {got[worst]!r}, oracle {orgb[worst]!r}), "
            f"uchar steps max {d8.max()} frac {np.mean(d8 > 0):.3g}")
    print(diag)
    assert rmse <= RMSE_MAX, diag
    assert np.all(err <= _channel_bound(orgb)), diag
    assert d8.max() <= 1 and np.mean(d8 > 0) <= U8_STEP_FRAC, diag
    gg.close()
    gc.close()


def test_shipped_full_canvas(ctx, scene, reference_maps, tmp_path):
    """The reference's 1024 x 768 frame: one-call and phased (batched gather)
    paths agree exactly, every pixel is written, and the PPM holds the canvas."""
    from pm355.image import read_ppm, write_ppm

    ds, g, c, *_ = reference_maps
    aabb = scene.aabb()
    gg = ctx.grid_build(g, aabb)
    gc = ctx.grid_build(c, aabb)
    W, H = RF["width"], RF["height"]
    cam = reference_camera(W, H)
    args = (RF["max_depth"], RF["k_global"], RF["r_global"], RF["k_caustic"], RF["r_caustic"])
    rgb, rgb8, hits = ctx.render_frame(ds, gg, gc, cam, *args)
    (brgb, brgb8, bhits), = ctx.render_frames_batched([(ds, gg, gc, cam, *args)])
    ctx.sync()
    assert hits == bhits and hits >= W * H // 2
    assert np.array_equal(_np(rgb).view(np.uint32), _np(brgb).view(np.uint32))
    assert np.array_equal(_np(rgb8), _np(brgb8))
    f = _np(rgb)
    assert np.all(np.isfinite(f)) and f.min() >= 0
    p = str(tmp_path / "frame.ppm")
    write_ppm(p, rgb8, W, H)
    img = read_ppm(p)
    assert np.array_equal(img[::-1].reshape(-1), _np(rgb8))
    # the canvas is not flat: walls of three colours and the light are seen
    assert len(np.unique(img.reshape(-1, 3), axis=0)) > 1000
    gg.close()
    gc.close()


def test_shipped_served_frames(ctx, scene):
    """The server renders the shipped scene with the reference's k and radii
    (batched gathers with k above 256) exactly as the one-frame-at-a-time
    calls do."""
    from test_gpu_serve import _one_at_a_time, _serve

    ds = ctx.upload_scene(scene, accel=True)
    spec = dict(dscene=ds, aabb=scene.aabb(), cell_size=0.0, n_global=200_000, n_caustic=20_000,
                camera=reference_camera(128, 96), frames=2, seed=77, max_depth=RF["max_depth"],
                k_global=RF["k_global"], r_global=RF["r_global"], k_caustic=RF["k_caustic"],
                r_caustic=RF["r_caustic"], max_paths=50_000_000)
    print("shipped scene served frames", flush=True)
    (out,), *_ = _serve(ctx, [spec])
    ref = _one_at_a_time(ctx, spec)
    assert np.array_equal(out, ref.reshape(out.shape))


def test_render_obj_cli(ctx, scene, tmp_path, capsys):
    """The headless renderer writes the canvas render_frame returns for the
    same scene, photon maps and parameters."""
    from pm355.image import read_ppm
    from pm355.render_obj import main

    obj = os.path.join(GOLDEN, "cornell_box_sphere", "CornellBox-Sphere.obj")
    out = str(tmp_path / "f.ppm")
    W, H = 96, 72
    assert main([obj, out, "--width", str(W), "--height", str(H), "--global-photons", "50000",
                 "--caustic-photons", "5000", "--seed", "5", "--pfm"]) == 0
    assert os.path.exists(str(tmp_path / "f.pfm"))
    ds = ctx.upload_scene(scene, accel=True)
    g, c, st = ctx.trace_photons(ds, 5, 50_000, 5_000, RF["max_bounces"])
    aabb = scene.aabb()
    gg, gc = ctx.grid_build(g, aabb), ctx.grid_build(c, aabb)
    _, rgb8, _ = ctx.render_frame(ds, gg, gc, reference_camera(W, H), RF["max_depth"], RF["k_global"],
                                  RF["r_global"], RF["k_caustic"], RF["r_caustic"])
    ctx.sync()
    assert np.array_equal(read_ppm(out)[::-1].reshape(-1), _np(rgb8))
    assert '"eye_ray_hits"' in capsys.readouterr().out
    gg.close()
    gc.close()
