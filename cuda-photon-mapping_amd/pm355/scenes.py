"""Synthetic scenes for the photon pipeline (host numpy arrays).

Scene arrays follow the pm355 C ABI (include/pm355.h):
  tris9    float32[T, 9]   triangle vertices v0.xyz v1.xyz v2.xyz
  normals9 float32[T, 9]   per-vertex normals n0 n1 n2
  models   PM_MODEL_DTYPE[M] (tri range, material, id)
A model is a contiguous range of triangles with one material, like the
reference's Model + Material (Model.cuh:24-30, Material.cuh:4-31).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PM_MODEL_DTYPE = np.dtype([
    ("tri_begin", "<u4"), ("tri_end", "<u4"),
    ("kd", "<f4", 3), ("ks", "<f4", 3), ("ke", "<f4", 3),
    ("refr_index", "<f4"), ("opaque", "<f4"), ("shininess", "<f4"),
    ("id", "<i4"), ("pad", "<i4"),
    ("aabb", "<f4", 6),
])

PM_CAMERA_DTYPE = np.dtype([
    ("position", "<f4", 3), ("forward", "<f4", 3), ("right", "<f4", 3), ("up", "<f4", 3),
    ("viewport_width", "<f4"), ("viewport_height", "<f4"), ("focal_length", "<f4"),
    ("width", "<i4"), ("height", "<i4"), ("pad", "<i4"),
])


@dataclass
class Scene:
    tris9: np.ndarray
    normals9: np.ndarray
    models: np.ndarray
    light_model: int
    light_intensity: tuple = (1.0, 1.0, 1.0)
    default_refr_index: float = 1.0
    names: list = field(default_factory=list)

    @property
    def tri_begin(self) -> np.ndarray:
        return np.concatenate([self.models["tri_begin"], self.models["tri_end"][-1:]]).astype(np.uint32)

    def aabb(self) -> tuple:
        """Scene AABB = union of model AABBs (Initializer.cu:197-200)."""
        v = self.tris9.reshape(-1, 3)
        return tuple(v.min(0).tolist()) + tuple(v.max(0).tolist())


def _quad(p0, p1, p2, p3, n):
    """Two triangles (p0 p1 p2), (p0 p2 p3) with a flat normal."""
    t = np.array([np.concatenate([p0, p1, p2]), np.concatenate([p0, p2, p3])], np.float32)
    nn = np.array([np.tile(n, 3), np.tile(n, 3)], np.float32)
    return t, nn


def uv_sphere(center, radius, slices=24, stacks=21):
    """Tessellated sphere with 2*slices*(stacks-1) triangles and smooth
    (radial) vertex normals; slices=24, stacks=21 gives 960 triangles.
    Triangle order: stack by stack, and within a stack slice by slice the
    upper triangle (a, b, d) then the lower one (d, b, c)."""
    c = np.asarray(center, np.float64)
    i = np.arange(stacks + 1, dtype=np.float64)[:, None]
    j = np.arange(slices + 1, dtype=np.float64)[None, :]
    th = np.pi * i / stacks
    ph = 2 * np.pi * j / slices
    d = np.stack(np.broadcast_arrays(np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)), -1)
    p = c + radius * d  # [stacks+1, slices+1, 3]
    tris, nrms = [], []
    for si in range(stacks):
        a, na = p[si, :-1], d[si, :-1]
        bb, nb = p[si + 1, :-1], d[si + 1, :-1]
        cc, nc = p[si + 1, 1:], d[si + 1, 1:]
        dd, nd = p[si, 1:], d[si, 1:]
        up = np.concatenate([a, bb, dd], 1), np.concatenate([na, nb, nd], 1)
        lo = np.concatenate([dd, bb, cc], 1), np.concatenate([nd, nb, nc], 1)
        if si == 0:
            t, n = lo
        elif si == stacks - 1:
            t, n = up
        else:
            t = np.stack([up[0], lo[0]], 1).reshape(-1, 9)
            n = np.stack([up[1], lo[1]], 1).reshape(-1, 9)
        tris.append(t)
        nrms.append(n)
    return np.concatenate(tris).astype(np.float32), np.concatenate(nrms).astype(np.float32)


def _model(tb, te, kd=(0, 0, 0), ks=(0, 0, 0), ke=(0, 0, 0), ior=1.0, opaque=1.0, shininess=1.0, mid=0):
    m = np.zeros((), PM_MODEL_DTYPE)
    m["tri_begin"], m["tri_end"] = tb, te
    m["kd"], m["ks"], m["ke"] = kd, ks, ke
    m["refr_index"], m["opaque"], m["shininess"] = ior, opaque, shininess
    m["id"] = mid
    return m


def cornell_box(sphere_slices=24, sphere_stacks=21) -> Scene:
    """The small Cornell-box scene of the benchmark configs: unit cube with 5
    diffuse walls (10 triangles; red and green side walls albedo 0.75, others
    white 0.75) and an open front (z = 1); a 0.25 x 0.25 square area light at
    the ceiling centre; a mirror sphere r = 0.2 and a glass sphere r = 0.2
    (IOR 1.5), 960 triangles each."""
    P = lambda *a: np.array(a, np.float32)
    parts = []  # (name, tris, normals, material kwargs)
    w = 0.75
    parts.append(("floor", *_quad(P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0), P(0, 1, 0)), dict(kd=(w, w, w))))
    parts.append(("ceiling", *_quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1), P(0, -1, 0)), dict(kd=(w, w, w))))
    parts.append(("backWall", *_quad(P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0), P(0, 0, 1)), dict(kd=(w, w, w))))
    parts.append(("leftWall", *_quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1), P(1, 0, 0)), dict(kd=(w, 0, 0))))
    parts.append(("rightWall", *_quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0), P(-1, 0, 0)), dict(kd=(0, w, 0))))
    ms, mn = uv_sphere((0.3, 0.2, 0.35), 0.2, sphere_slices, sphere_stacks)
    parts.append(("mirrorSphere", ms, mn, dict(ks=(1, 1, 1), shininess=0.0)))
    gs, gn = uv_sphere((0.7, 0.2, 0.62), 0.2, sphere_slices, sphere_stacks)
    parts.append(("glassSphere", gs, gn, dict(ks=(1, 1, 1), opaque=0.0, ior=1.5)))
    y = np.float32(0.999)
    lo, hi = np.float32(0.375), np.float32(0.625)
    parts.append(("light", *_quad(P(lo, y, lo), P(hi, y, lo), P(hi, y, hi), P(lo, y, hi), P(0, -1, 0)),
                  dict(kd=(0.78, 0.78, 0.78), ke=(1, 1, 1))))
    tris, nrms, models, names = [], [], [], []
    t0 = 0
    for i, (name, t, n, mat) in enumerate(parts):
        tris.append(t)
        nrms.append(n)
        models.append(_model(t0, t0 + len(t), mid=i, **mat))
        names.append(name)
        t0 += len(t)
    models = np.array(models, PM_MODEL_DTYPE)
    return Scene(np.concatenate(tris), np.concatenate(nrms), models, light_model=names.index("light"), names=names)


def camera(position, direction, up, vfov_deg, width, height) -> np.ndarray:
    """Camera(pos, dir, up, fov, W, H) (Camera.cuh:43-66) evaluated in float32
    in the reference's host order."""
    f32 = np.float32
    pos = np.asarray(position, f32)
    d = np.asarray(direction, f32)
    u = np.asarray(up, f32)
    aspect = f32(width) / f32(height)
    theta = f32(f32(vfov_deg) * f32(3.141592653589793)) / f32(180.0)
    # tan in double, rounded once (tanf implementations differ by an ulp)
    h = f32(math.tan(float(theta / f32(2))))
    vh = f32(2.0) * h
    vw = aspect * vh

    def norm(v):
        if not np.any(v):
            return v
        ln = f32(np.sqrt(f32(f32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))
        return (v / ln).astype(f32)

    def cross(a, b):
        return np.array([a[1] * b[2] - a[2] * b[1], -(a[0] * b[2] - a[2] * b[0]), a[0] * b[1] - a[1] * b[0]], f32)

    fwd = norm(d)
    right = norm(cross(fwd, u))
    upv = cross(right, fwd)
    c = np.zeros((), PM_CAMERA_DTYPE)
    c["position"], c["forward"], c["right"], c["up"] = pos, fwd, right, upv
    c["viewport_width"], c["viewport_height"], c["focal_length"] = vw, vh, 1.0
    c["width"], c["height"] = width, height
    return c


def cornell_camera(width=320, height=240) -> np.ndarray:
    """The reference camera (kernel.cu:29-30: (-0.01, 0.795, 2.35) looking
    down -z, 60 degree vertical field of view) mapped to the unit box."""
    return camera((0.5, 0.5, 1.85), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 60.0, width, height)


def reference_camera(width=1024, height=768) -> np.ndarray:
    """The reference's camera for its CornellBox-Sphere scene, at its default
    1024 x 768 canvas (kernel.cu:14-15, 29-30)."""
    return camera((-0.00999999046, 0.795000017, 2.35000001), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 60.0,
                  width, height)


# The reference's frame on its shipped scene: photon counts
# (Initializer.cu:145-146), photon bounces (PHOTON_BOUNCE_NUM,
# PhotonTracing.cu:20), gather k (kernel.cu:25-26), gather radii
# (Raytracer.cuh:187, 200), eye-ray depth (kernel.cu:23), canvas
# (kernel.cu:14-15).
REFERENCE_FRAME = dict(n_global=1_000_000, n_caustic=100_000, max_bounces=10, max_depth=4, k_global=2000,
                       r_global=0.1, k_caustic=400, r_caustic=0.2, width=1024, height=768)


def random_scene(seed: int, n_spheres=None, tri_range=(200, 20000)) -> Scene:
    """A scene of the batched-workload distribution: the open unit box of
    cornell_box (5 walls and the ceiling light) plus 1-8 spheres, each
    tessellated to 200-20k triangles (log-uniform), radius 0.05-0.2, with a
    material drawn from diffuse (albedo U[0.2, 0.9]), mirror, or glass
    (IOR U[1.3, 1.7]). Spheres may overlap each other and the walls."""
    rng = np.random.default_rng(seed)
    base = cornell_box()
    keep = [i for i, nm in enumerate(base.names) if nm not in ("mirrorSphere", "glassSphere")]
    tris, nrms, mats, names = [], [], [], []
    for i in keep:
        m = base.models[i]
        tris.append(base.tris9[m["tri_begin"]:m["tri_end"]])
        nrms.append(base.normals9[m["tri_begin"]:m["tri_end"]])
        mats.append(dict(kd=tuple(m["kd"]), ks=tuple(m["ks"]), ke=tuple(m["ke"]), ior=float(m["refr_index"]),
                         opaque=float(m["opaque"]), shininess=float(m["shininess"])))
        names.append(base.names[i])
    ns = int(rng.integers(1, 9)) if n_spheres is None else int(n_spheres)
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    for s_i in range(ns):
        target = float(np.exp(rng.uniform(lo, hi)))
        slices = max(3, int(round(np.sqrt(target / 1.8))))
        stacks = max(2, int(round(target / (2 * slices))) + 1)
        r = float(rng.uniform(0.05, 0.2))
        c = (float(rng.uniform(r, 1 - r)), float(rng.uniform(r, 0.9 - r)), float(rng.uniform(r, 1 - r)))
        t, n = uv_sphere(c, r, slices, stacks)
        kind = int(rng.integers(0, 3))
        if kind == 0:
            a = float(rng.uniform(0.2, 0.9))
            mat = dict(kd=(a, a * float(rng.uniform(0.5, 1)), a * float(rng.uniform(0.5, 1))))
        elif kind == 1:
            mat = dict(ks=(1, 1, 1), shininess=0.0)
        else:
            mat = dict(ks=(1, 1, 1), opaque=0.0, ior=float(rng.uniform(1.3, 1.7)))
        tris.append(t)
        nrms.append(n)
        mats.append(mat)
        names.append(f"sphere{s_i}")
    models, t0 = [], 0
    for i, (t, mat) in enumerate(zip(tris, mats)):
        models.append(_model(t0, t0 + len(t), mid=i, **mat))
        t0 += len(t)
    return Scene(np.concatenate(tris), np.concatenate(nrms), np.array(models, PM_MODEL_DTYPE),
                 light_model=names.index("light"), names=names)
