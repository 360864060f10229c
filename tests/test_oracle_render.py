"""CPU tests pinning the oracle's frame renderer (render_trace semantics).

* Emission-only frame (photons far away, so every estimate is empty): the
  image is HDR(emission) along the ray tree, so light pixels and pixels that
  see the light in the mirror take the closed-form value
  (1 - e^{-1/0.6})^{1/2.2} * 255 -> 231, everything else 0; the set of direct
  light pixels matches an independent float64 ray-plane test.
* A single diffuse hit composes the pinned gather + radiance oracles exactly
  as Raytracer.cuh:161-211 does.
"""
import numpy as np
import pytest

import pyoracle as O
from pm355.scenes import cornell_box, cornell_camera


@pytest.fixture(scope="module")
def scene():
    return cornell_box()


def _far_grid(aabb):
    ph = np.zeros((9, 4), np.float32)
    ph[0:3] = 50.0  # outside every gather radius
    ph[3:6] = 1.0
    ph[6:9, :] = np.array([[0], [-1], [0]], np.float32)
    return O.OracleGrid(ph, (0, 0, 0, 60, 60, 60), cell_size=1.0)


def _primary_dirs(cam):
    W, H = int(cam["width"].item()), int(cam["height"].item())
    f32 = np.float32
    i, j = np.meshgrid(np.arange(W), np.arange(H))
    u = ((i + f32(0.5)) / f32(W)).astype(f32)
    v = ((j + f32(0.5)) / f32(H)).astype(f32)
    pos, fwd = cam["position"].astype(np.float64), cam["forward"].astype(np.float64)
    hor = cam["right"].astype(np.float64) * cam["viewport_width"]
    ver = cam["up"].astype(np.float64) * cam["viewport_height"]
    ll = (pos + fwd) * cam["focal_length"] - hor / 2 - ver / 2
    d = ll[None, None] + u[..., None] * hor + v[..., None] * ver - pos
    return pos, d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_emission_only_frame(scene):
    cam = cornell_camera(160, 120)
    far = _far_grid(scene.aabb())
    rgb, rgb8, hits = O.render(scene, far, far, cam, 4, 8, 0.05, 8, 0.05)
    img = rgb8.reshape(120, 160, 3)
    lit = np.float32(1.0) - np.exp(np.float32(-1.0) / np.float32(0.6), dtype=np.float32)
    expect = int(np.float32(lit ** np.float32(1 / 2.2)) * 255)
    assert expect == 231
    vals = np.unique(img)
    assert set(vals.tolist()) <= {0, expect}
    # direct view of the light quad (y = 0.999, x, z in [0.375, 0.625])
    pos, d = _primary_dirs(cam)
    t = (0.999 - pos[1]) / d[..., 1]
    p = pos + t[..., None] * d
    direct = (t > 0) & (np.abs(p[..., 0] - 0.5) < 0.125) & (np.abs(p[..., 2] - 0.5) < 0.125)
    # pixels within one pixel of the quad's edge may go either way
    sure = direct & (np.abs(p[..., 0] - 0.5) < 0.12) & (np.abs(p[..., 2] - 0.5) < 0.12)
    assert np.all(img[sure] == expect)
    assert sure.sum() > 50
    # the mirror sphere shows the light as well
    assert (img[..., 0] == expect).sum() > direct.sum()


def test_single_diffuse_hit_composition(scene):
    g, c, _ = O.trace_photons(scene, 3, 30000, 3000, 8)
    aabb = scene.aabb()
    og = O.OracleGrid(g, aabb, cell_size=0.05)
    oc = O.OracleGrid(c, aabb, cell_size=0.05)
    cam = cornell_camera(64, 48)
    rgb, rgb8, _ = O.render(scene, og, oc, cam, 4, 50, 0.05, 20, 0.05)
    pos, d = _primary_dirs(cam)
    # pick pixels whose primary ray hits the back wall (diffuse, opaque, Ks = 0)
    W = 64
    checked = 0
    for (jj, ii) in [(j, i) for j in range(4, 48, 6) for i in range(4, 64, 7)]:
        if checked >= 8:
            break
        dd = d[jj, ii].astype(np.float32)
        # the oracle's own primary ray (float32, reference order) -> hit point
        t, m, n = O.scene_hit(scene, np.array([[*cam["position"]], [*dd]], np.float32).reshape(6, 1))
        if m[0] != 2:  # backWall
            continue
        ip = (cam["position"] + dd * t[0]).astype(np.float32).reshape(3, 1)
        gi, gd, gc = og.knn(ip, 0.05, 50)
        lg, _ = O.radiance(gi, gd, gc, og.photons, n, 50)
        ci, cd, cc = oc.knn(ip, 0.05, 20)
        lc = O.radiance(ci, cd, cc, oc.photons, n, 20)[0] if cc[0] == 20 else np.zeros((3, 1), np.float32)
        cur = np.minimum(lg[:, 0] * np.float32(0.75) * np.float32(4.0) + lc[:, 0], np.float32(1.0))
        mapped = np.minimum((np.float32(1) - np.exp(-np.float32(1 / 0.6) * cur)) ** np.float32(1 / 2.2), 1.0)
        got = rgb[:, jj * W + ii]
        assert np.allclose(got, mapped, atol=2e-3), (got, mapped)
        checked += 1
    assert checked >= 2
