"""CPU tests pinning the oracle's frame renderer (render_trace semantics).

* Emission-only frame (photons far away, so every estimate is empty): the
  image is HDR(emission) along the ray tree, so light pixels and pixels that
  see the light in the mirror take the closed-form value
  (1 - e^{-1/0.6})^{1/2.2} * 255 -> 231, everything else 0; the set of direct
  light pixels matches an independent float64 ray-plane test.
* A single diffuse hit composes the pinned gather + radiance oracles exactly
  as Raytracer.cuh:161-211 does.
* A tilted mirror reflecting onto an emissive wall: every pixel equals
  HDR(E * Ks * opaque * cos(2 theta)^n) from float64 geometry (law of
  reflection, Raytracer.cuh:228-235).
* A glass quad and a glass slab in front of an emissive patch: the lit
  pixels are those whose float64 Snell path (with the 0.01 origin offsets of
  Ray.cuh:85-96) lands on the patch, with value HDR(E (1 - opaque)^m); a
  straight ray or a missing / reversed offset lights other pixels.
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


# ---------------------------------------------------------------- closed-form mirror and refraction frames

from pm355.scenes import PM_MODEL_DTYPE, Scene, _model, _quad, camera as make_camera  # noqa: E402


def _tilted_quad(c, n, half):
    c, n = np.asarray(c, np.float64), np.asarray(n, np.float64)
    v = np.array([0.0, 1.0, 0.0])
    u = np.cross(v, n)
    P = lambda a, b: (c + half * (a * u + b * v)).astype(np.float32)
    t, nn = _quad(P(-1, -1), P(1, -1), P(1, 1), P(-1, 1), n.astype(np.float32))
    return t, nn, u, v


def _scene_of(parts, light):
    tris, nrms, models, t0 = [], [], [], 0
    for i, (t, n, mat) in enumerate(parts):
        tris.append(t)
        nrms.append(n)
        models.append(_model(t0, t0 + len(t), mid=i, **mat))
        t0 += len(t)
    return Scene(np.concatenate(tris), np.concatenate(nrms), np.array(models, PM_MODEL_DTYPE), light_model=light)


def _plane_hit(o, d, p0, n):
    t = ((np.asarray(p0) - o) * n).sum(-1) / (d * n).sum(-1)
    return t, o + t[..., None] * d


def _hdr(x):
    return (1 - np.exp(-x / 0.6)) ** (1 / 2.2)


@pytest.mark.parametrize("ks,opaque,shin", [(0.7, 1.0, 3.0), (0.5, 0.6, 1.0)])
def test_mirror_bounce_coefficient(ks, opaque, shin):
    """Raytracer.cuh:228-235: a mirror hit multiplies the ray coefficient by
    Ks * opaque * max(dot(refl, -in), 0)^shininess, and dot(refl, -in) =
    cos(2 theta) for incidence angle theta (law of reflection). The reflected
    ray lands on an emissive wall (Ke = E <= 1, clamped at 1), so the pixel
    is HDR(E * Ks * opaque * cos(2 theta)^n) exactly."""
    E = 0.8
    alpha = np.radians(25)
    nm = np.array([np.sin(alpha), 0, np.cos(alpha)])
    mt, mn, u, v = _tilted_quad((0.5, 0.5, 0.0), nm, 0.6)
    wt, wn = _quad(*(np.array(p, np.float32) for p in ([1.5, -2, -2], [1.5, 3, -2], [1.5, 3, 4], [1.5, -2, 4])),
                   np.array([-1, 0, 0], np.float32))
    sc = _scene_of([(mt, mn, dict(ks=(ks, ks, ks), opaque=opaque, shininess=shin, ior=1.5)),
                    (wt, wn, dict(ke=(E, E, E)))], light=1)
    W, H = 160, 120
    cam = make_camera((0.5, 0.5, 2.0), (0, 0, -1), (0, 1, 0), 40.0, W, H)
    far = _far_grid(sc.aabb())
    rgb, _, _ = O.render(sc, far, far, cam, 4, 8, 0.05, 8, 0.05)
    pos, d = _primary_dirs(cam)
    t, p = _plane_hit(pos, d, (0.5, 0.5, 0.0), nm)
    on = (np.abs(((p - [0.5, 0.5, 0]) * u).sum(-1)) < 0.58) & (np.abs(((p - [0.5, 0.5, 0]) * v).sum(-1)) < 0.58)
    cos_t = -(d * nm).sum(-1)
    refl = d + 2 * cos_t[..., None] * nm
    _, q = _plane_hit(p + 0.01 * nm, refl, (1.5, 0, 0), np.array([1.0, 0, 0]))
    lands = (q[..., 1] > -1.9) & (q[..., 1] < 2.9) & (q[..., 2] > -1.9) & (q[..., 2] < 3.9)
    sel = on & lands & (t > 0)
    assert sel.sum() > 2000
    cos2 = 2 * cos_t ** 2 - 1
    want = _hdr(E * ks * opaque * np.maximum(cos2, 0) ** shin)
    got = rgb[0].reshape(H, W)
    assert np.allclose(got[sel], want[sel], atol=3e-6, rtol=1e-5), np.abs(got[sel] - want[sel]).max()
    # the incidence angles span a real range, so cos(theta)^n would not pass
    wrong = _hdr(E * ks * opaque * cos_t ** shin)
    assert np.abs(wrong[sel] - want[sel]).min() > 1e-3


def _snell(d, n, n1, n2):
    """Refraction of unit d at unit normal n (facing against d) by the
    vector form of Snell's law, in float64."""
    c1 = -(d * n).sum(-1, keepdims=True)
    eta = n1 / n2
    k = 1 - eta ** 2 * (1 - c1 ** 2)
    return eta * d + (eta * c1 - np.sqrt(np.maximum(k, 0))) * n


@pytest.mark.parametrize("slab", [False, True])
def test_refraction_through_glass_onto_emitter(slab):
    """A glass quad (or a slab: two parallel faces of one model) with
    opaque 0.25 in front of an emissive patch. The deferred refraction ray
    (RaytracePlanner::push_refraction, Raytracer.cuh:213-226) carries
    coefficient (1 - opaque) per interface, bends by Snell's law (n 1 -> 1.5,
    and 1.5 -> 1 leaving the slab) and starts 0.01 behind the surface
    (Ray.cuh:85-96). Pixels are HDR(E (1 - opaque)^m) where the float64
    refracted path lands on the patch and 0 where it misses; a straight ray,
    an unrefracted exit or an origin without the 0.01 offset would light
    other pixels."""
    E, op = 0.9, 0.25
    beta = np.radians(50)
    ng = np.array([np.sin(beta), 0, np.cos(beta)])
    c0 = np.array([0.5, 0.5, 0.0])
    ft, fn, u, v = _tilted_quad(c0, ng, 0.7)
    glass = dict(opaque=op, ior=1.5)
    thick = 0.15
    if slab:
        bt, bn, _, _ = _tilted_quad(c0 - thick * ng, -ng, 0.7)
        gt, gn = np.concatenate([ft, bt]), np.concatenate([fn, bn])
    else:
        gt, gn = ft, fn
    P = lambda *a: np.array(a, np.float32)
    et, en = _quad(P(0.35, 0.3, -0.6), P(0.65, 0.3, -0.6), P(0.65, 0.7, -0.6), P(0.35, 0.7, -0.6), P(0, 0, 1))
    sc = _scene_of([(gt, gn, glass), (et, en, dict(ke=(E, E, E)))], light=1)
    W, H = 400, 300
    cam = make_camera((0.5, 0.5, 2.0), (0, 0, -1), (0, 1, 0), 20.0, W, H)
    far = _far_grid(sc.aabb())
    rgb, _, _ = O.render(sc, far, far, cam, 4, 8, 0.05, 8, 0.05)
    got = rgb[0].reshape(H, W)
    pos, d = _primary_dirs(cam)

    def land(bend=True, offset=0.01):
        t, p = _plane_hit(pos, d, c0, ng)
        r = _snell(d, ng, 1.0, 1.5) if bend else d
        o = p - offset * ng
        if slab:
            t2, p2 = _plane_hit(o, r, c0 - thick * ng, ng)
            r = _snell(r, ng, 1.5, 1.0) if bend else r
            o = p2 - offset * ng
        _, q = _plane_hit(o, r, (0, 0, -0.6), np.array([0, 0, 1.0]))
        on = (np.abs(((p - c0) * u).sum(-1)) < 0.68) & (np.abs(((p - c0) * v).sum(-1)) < 0.68) & (t > 0)
        gx, gy = np.minimum(q[..., 0] - 0.35, 0.65 - q[..., 0]), np.minimum(q[..., 1] - 0.3, 0.7 - q[..., 1])
        inside = np.minimum(gx, gy)
        return on, inside

    on, inside = land()
    sure = on & (np.abs(inside) > 5e-4)
    lit = sure & (inside > 0)
    m = 2 if slab else 1
    assert lit.sum() > 500 and (sure & ~lit).sum() > 500
    assert np.allclose(got[lit], _hdr(E * (1 - op) ** m), atol=3e-6)
    assert np.all(got[sure & ~lit] == 0)
    for alt in (dict(bend=False), dict(offset=0.0), dict(offset=-0.01)):
        on_a, inside_a = land(**alt)
        both = sure & on_a & (np.abs(inside_a) > 5e-4)
        assert ((inside_a > 0) != (inside > 0))[both].sum() > 20, alt
