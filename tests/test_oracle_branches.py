"""CPU tests pinning the oracle tracer's stochastic branches with closed forms.

Each test traces >= 1e5 light paths through a scene whose branch
probabilities are known in closed form and compares observed frequencies
with them, within 5 sigma of the binomial (or Kolmogorov) spread. Every
check is chosen so that a plausible restatement mistake fails it:

* Russian roulette draws e in [0, 2) (PhotonTracing.cu:259): in a closed box
  of one material (Kd, Ks) and grey photon power, a bounce is diffuse with
  probability max(Kd)/2, specular with max(Ks)/2, absorbed otherwise
  (:261-275); diffuse and absorbed-with-Ks != 1 store (:308-313). With Ks = 0
  the photons per path follow a truncated geometric law of ratio max(Kd)/2.
* A Ks = 1 mirror continues with probability 1/2 and never stores; its ray
  obeys the law of reflection and leaves from p + 0.01 n (Ray.cuh:64-71).
* A glass sphere (IOR 1.5): refraction happens with probability
  1 - Schlick(cos) (PhotonTracing.cu:215-243), never beyond the critical
  angle, and the refracted ray obeys Snell's law, leaving from p - 0.01 n
  (Ray.cuh:85-96). Entering / leaving is decided from the geometry here,
  independently of the oracle's medium stack.
* Diffuse bounces are uniform on the hemisphere (Ray.cuh:16-56: Marsaglia
  sphere points folded by rejection), so cos theta ~ U(0, 1).
* Emission (PhotonTracing.cu:345-367, Model.cuh:202-243): triangle by
  roundf(u * (tris - 1)), barycentrics x = u1, y = u2 (1 - x); directions
  normalize(U[-1,1]^3) folded into the light's hemisphere and accepted with
  probability cos theta, i.e. density ~ cos theta / max|d_i|^3.
"""
import numpy as np
import pytest

import pyoracle as O
from pm355.scenes import PM_MODEL_DTYPE, Scene, _model, _quad, uv_sphere

N_PATHS = 100_000
Z5 = 5.0


def _box_parts(wall_mat, light_mat=None):
    P = lambda *a: np.array(a, np.float32)
    walls = [
        _quad(P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0), P(0, 1, 0)),    # floor
        _quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1), P(0, -1, 0)),   # ceiling
        _quad(P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0), P(0, 0, 1)),    # back
        _quad(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1), P(1, 0, 0)),    # left
        _quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0), P(-1, 0, 0)),   # right
        _quad(P(0, 0, 1), P(0, 1, 1), P(1, 1, 1), P(1, 0, 1), P(0, 0, -1)),   # front (closed)
    ]
    parts = [(f"wall{i}", t, n, wall_mat) for i, (t, n) in enumerate(walls)]
    y, lo, hi = np.float32(0.999), np.float32(0.375), np.float32(0.625)
    lt, ln = _quad(P(lo, y, lo), P(hi, y, lo), P(hi, y, hi), P(lo, y, hi), P(0, -1, 0))
    parts.append(("light", lt, ln, light_mat if light_mat is not None else wall_mat))
    return parts


def _scene(parts):
    tris, nrms, models, names, t0 = [], [], [], [], 0
    for i, (name, t, n, mat) in enumerate(parts):
        tris.append(t)
        nrms.append(n)
        models.append(_model(t0, t0 + len(t), mid=i, **mat))
        names.append(name)
        t0 += len(t)
    return Scene(np.concatenate(tris), np.concatenate(nrms), np.array(models, PM_MODEL_DTYPE),
                 light_model=names.index("light"), names=names)


def closed_box(kd, ks=(0, 0, 0)):
    return _scene(_box_parts(dict(kd=kd, ks=ks)))


SPHERE_C, SPHERE_R = (0.5, 0.42, 0.5), 0.3


def box_with_sphere(sphere_mat):
    parts = _box_parts(dict(kd=(0.6, 0.6, 0.6)))
    st, sn = uv_sphere(SPHERE_C, SPHERE_R, 24, 21)
    parts.insert(6, ("sphere", st, sn, sphere_mat))
    return _scene(parts)


CUBE_C, CUBE_H = np.array([0.5, 0.4, 0.5]), 0.25


def box_with_cube(cube_mat):
    """A glass cube (flat outward normals) inside the closed box: inside a
    sphere every chord meets the surface at the entry angle, so only flat
    faces at right angles produce total internal reflection."""
    c, h = CUBE_C, CUBE_H
    P = lambda *a: (c + h * np.array(a, np.float64)).astype(np.float32)
    N = lambda *a: np.array(a, np.float32)
    faces = [
        _quad(P(-1, -1, -1), P(1, -1, -1), P(1, -1, 1), P(-1, -1, 1), N(0, -1, 0)),
        _quad(P(-1, 1, -1), P(-1, 1, 1), P(1, 1, 1), P(1, 1, -1), N(0, 1, 0)),
        _quad(P(-1, -1, -1), P(-1, 1, -1), P(1, 1, -1), P(1, -1, -1), N(0, 0, -1)),
        _quad(P(-1, -1, 1), P(1, -1, 1), P(1, 1, 1), P(-1, 1, 1), N(0, 0, 1)),
        _quad(P(-1, -1, -1), P(-1, -1, 1), P(-1, 1, 1), P(-1, 1, -1), N(-1, 0, 0)),
        _quad(P(1, -1, -1), P(1, 1, -1), P(1, 1, 1), P(1, -1, 1), N(1, 0, 0)),
    ]
    parts = _box_parts(dict(kd=(0.6, 0.6, 0.6)))
    parts.insert(6, ("cube", np.concatenate([f[0] for f in faces]), np.concatenate([f[1] for f in faces]), cube_mat))
    return _scene(parts)


def _binom_ok(k, n, p, z=Z5):
    sd = np.sqrt(n * p * (1 - p))
    return abs(k - n * p) <= z * max(sd, 1e-12), (k, n * p, sd)


def _ks_uniform(x):
    """Kolmogorov distance of a sample from U(0, 1) and the 1e-6 critical
    value sqrt(ln(2/alpha) / (2n))."""
    x = np.sort(np.asarray(x, np.float64))
    n = x.size
    i = np.arange(1, n + 1)
    d = max(np.max(i / n - x), np.max(x - (i - 1) / n))
    return d, np.sqrt(np.log(2 / 1e-6) / (2 * n))


def _per_path_counts(ev, n_paths):
    return np.bincount(ev["path"][ev["stored"] > 0], minlength=n_paths)


# ---------------------------------------------------------------- Russian roulette


@pytest.mark.parametrize("kd", [(0.6, 0.6, 0.6), (0.2, 0.7, 0.4)])
def test_russian_roulette_diffuse_box(kd):
    """Closed box, one material, Ks = 0: every bounce stores; it continues
    (diffuse) with probability q = max(Kd) / 2; photons per path K follow
    P(K = j) = (1 - q) q^(j-1) for j < B and q^(B-1) for j = B."""
    B = 8
    sc = closed_box(kd)
    ev = O.trace_events(sc, 11, N_PATHS, B)
    q = max(kd) / 2
    assert set(np.unique(ev["dest"]).tolist()) <= {O.EV_DIFFUSE, O.EV_ABSORB}
    assert np.all(ev["stored"] == 1)  # every bounce stores, never in the caustic map
    ok, info = _binom_ok(int((ev["dest"] == O.EV_DIFFUSE).sum()), len(ev), q)
    assert ok, info
    K = _per_path_counts(ev, N_PATHS)
    assert K.min() >= 1 and K.max() <= B  # the box is closed: every emitted photon lands
    for j in range(1, B + 1):
        pj = (1 - q) * q ** (j - 1) if j < B else q ** (B - 1)
        ok, info = _binom_ok(int((K == j).sum()), N_PATHS, pj)
        assert ok, (j, info)
    # the same statistic end to end through the map builder
    _, _, st = O.trace_photons(sc, 11, 10 * N_PATHS, 1, B, max_paths=N_PATHS)
    assert st[0] == N_PATHS and st[1] == K.sum() and st[2] == 0


def test_russian_roulette_three_way_split():
    """Kd = 0.3 grey, Ks = (0.5, 0.2, 0.1): diffuse 0.15, specular 0.25,
    absorb 0.6 (e = 2u against pd = 0.3 and pd + ps = 0.8); diffuse and
    absorb store (Ks != 1), specular does not."""
    ev = O.trace_events(closed_box((0.3, 0.3, 0.3), (0.5, 0.2, 0.1)), 12, N_PATHS, 8)
    n = len(ev)
    for dest, p in ((O.EV_DIFFUSE, 0.15), (O.EV_SPECULAR, 0.25), (O.EV_ABSORB, 0.6)):
        ok, info = _binom_ok(int((ev["dest"] == dest).sum()), n, p)
        assert ok, (dest, info)
    assert np.all((ev["stored"] > 0) == (ev["dest"] != O.EV_SPECULAR))


# ---------------------------------------------------------------- mirror


def test_mirror_continues_half_and_never_stores():
    ev = O.trace_events(box_with_sphere(dict(ks=(1, 1, 1), shininess=0.0)), 13, N_PATHS, 10)
    m = ev[ev["model"] == 6]
    assert len(m) > 20_000
    assert not np.any(m["dest"] == O.EV_DIFFUSE) and not np.any(m["stored"])
    ok, info = _binom_ok(int((m["dest"] == O.EV_SPECULAR).sum()), len(m), 0.5)
    assert ok, info
    s = m[m["dest"] == O.EV_SPECULAR]
    d, n, nd = (s[f].astype(np.float64) for f in ("d", "n", "next_d"))
    # law of reflection: equal angles, coplanar, unit; origin p + 0.01 n
    assert np.allclose((nd * n).sum(1), -(d * n).sum(1), atol=2e-6)
    assert np.allclose((nd * np.cross(d, n)).sum(1), 0, atol=2e-6)
    assert np.allclose(np.linalg.norm(nd, axis=1), 1, atol=2e-6)
    assert np.allclose(s["next_o"], s["p"].astype(np.float64) + 0.01 * n, atol=1e-6)


def test_map_choice_follows_path_history():
    """PathOperator.cuh:28-37: a photon goes to the caustic map iff the
    previous event was specular or refractive and no diffuse bounce came
    before; and the stored events are exactly the maps' photons in order."""
    sc = box_with_sphere(dict(ks=(1, 1, 1), shininess=0.0))
    n_paths = 20_000
    ev = O.trace_events(sc, 14, n_paths, 10)
    first = np.r_[True, ev["path"][1:] != ev["path"][:-1]]
    prev = np.r_[-1, ev["dest"][:-1]]
    prev[first] = -1
    dif_before = np.zeros(len(ev), bool)
    seen = False
    for i in range(len(ev)):
        if first[i]:
            seen = False
        dif_before[i] = seen
        seen = seen or ev["dest"][i] == O.EV_DIFFUSE
    stored = ev["stored"] > 0
    want_caustic = np.isin(prev, (O.EV_SPECULAR, O.EV_REFRACT)) & ~dif_before
    assert np.array_equal(ev["stored"][stored] == 2, want_caustic[stored])
    assert (ev["stored"] == 2).sum() > 1000
    g, c, st = O.trace_photons(sc, 14, 10 * n_paths, 10 * n_paths, 10, max_paths=n_paths)
    for kind, m in ((1, g), (2, c)):
        e = ev[ev["stored"] == kind]
        assert m.shape[1] >= len(e) and np.array_equal(m[0:3, :len(e)].T, e["p"])
        assert np.array_equal(m[6:9, :len(e)].T, e["d"])


# ---------------------------------------------------------------- glass


@pytest.fixture(scope="module")
def glass_events():
    sc = box_with_cube(dict(ks=(1, 1, 1), opaque=0.0, ior=1.5))
    ev = O.trace_events(sc, 15, N_PATHS, 10)
    g = ev[ev["model"] == 6]
    # entering / leaving from the geometry: rays that start inside the cube
    # (refracted or reflected inward, origin 0.01 below a face) leave it
    inside = np.all(np.abs(g["o"].astype(np.float64) - CUBE_C) < CUBE_H, axis=1)
    entering = ~inside
    n1 = np.where(entering, 1.0, 1.5)
    n2 = np.where(entering, 1.5, 1.0)
    # the oracle's medium stack agrees with the geometry
    agree = (g["n1"] == n1.astype(np.float32)) & (g["n2"] == n2.astype(np.float32))
    assert agree.mean() > 0.999
    g, n1, n2 = g[agree], n1[agree], n2[agree]
    cos = -(g["d"].astype(np.float64) * g["n"].astype(np.float64)).sum(1)
    return g, n1, n2, cos


def test_glass_total_internal_reflection(glass_events):
    g, n1, n2, cos = glass_events
    cos_c = np.sqrt(1 - (1.0 / 1.5) ** 2)  # cos(asin(n2 / n1)) leaving the glass
    leaving = n1 > n2
    tir = leaving & (cos < cos_c - 1e-5)
    free = ~leaving | (cos > cos_c + 1e-5)
    assert tir.sum() > 2000
    assert np.all(g["flags"][tir] & 1) and not np.any(g["dest"][tir] == O.EV_REFRACT)
    assert not np.any(g["flags"][free] & 1)


def test_glass_fresnel_schlick(glass_events):
    """P(refract) = 1 - (R0 + (1 - R0)(1 - cos)^5), R0 = ((n1-n2)/(n1+n2))^2:
    0.96 at normal incidence for IOR 1.5. Checked per cos bin, entering and
    leaving separately."""
    g, n1, n2, cos = glass_events
    free = ~(g["flags"] & 1).astype(bool)
    r0 = ((n1 - n2) / (n1 + n2)) ** 2
    assert np.isclose(r0[0], 0.04)
    p = 1 - (r0 + (1 - r0) * (1 - cos) ** 5)
    refr = g["dest"] == O.EV_REFRACT
    checked = 0
    for side in (n1 < n2, n1 > n2):
        sel = free & side
        edges = np.quantile(cos[sel], np.linspace(0, 1, 11))
        for a, b in zip(edges[:-1], edges[1:]):
            s = sel & (cos >= a) & (cos <= b)
            exp, sd = p[s].sum(), np.sqrt((p[s] * (1 - p[s])).sum())
            assert abs(refr[s].sum() - exp) <= Z5 * sd + 1, (a, b, refr[s].sum(), exp, sd)
            checked += s.sum()
    assert checked > 30_000
    # the normal-incidence bin: close to 0.96
    ni = free & (cos > 0.99)
    assert ni.sum() > 500 and abs(refr[ni].mean() - 0.96) < 0.03
    # when not refracted, Russian roulette on Kd = 0, Ks = 1: specular half the time, never stored
    nr = ~refr
    ok, info = _binom_ok(int((g["dest"][nr] == O.EV_SPECULAR).sum()), int(nr.sum()), 0.5)
    assert ok, info
    assert not np.any(g["stored"])


def test_glass_refraction_obeys_snell(glass_events):
    g, n1, n2, cos = glass_events
    s = (g["dest"] == O.EV_REFRACT) & ~(g["flags"] & 2).astype(bool)
    d, n, t = (g[f][s].astype(np.float64) for f in ("d", "n", "next_d"))
    sin_i = np.linalg.norm(np.cross(d, n), axis=1)
    sin_t = np.linalg.norm(np.cross(t, n), axis=1)
    assert s.sum() > 20_000
    assert np.allclose(n1[s] * sin_i, n2[s] * sin_t, atol=2e-5)
    assert np.all((t * n).sum(1) < 0)  # transmitted through the surface
    assert np.allclose((t * np.cross(d, n)).sum(1), 0, atol=2e-5)  # plane of incidence
    assert np.allclose(np.linalg.norm(t, axis=1), 1, atol=2e-5)
    assert np.allclose(g["next_o"][s], g["p"][s].astype(np.float64) - 0.01 * n, atol=1e-6)


# ---------------------------------------------------------------- diffuse bounce and emission


def test_diffuse_bounce_uniform_on_hemisphere():
    ev = O.trace_events(closed_box((0.9, 0.9, 0.9)), 16, N_PATHS // 2, 8)
    s = ev[ev["dest"] == O.EV_DIFFUSE]
    n, nd = s["n"].astype(np.float64), s["next_d"].astype(np.float64)
    c = (n * nd).sum(1)
    assert len(s) > 40_000 and c.min() >= 0
    d, crit = _ks_uniform(c)
    assert d < crit, (d, crit)  # a cosine-weighted bounce would give D = 0.25
    # azimuth on the floor (n = +y): uniform in (-pi, pi]
    fl = np.abs(n[:, 1] - 1) < 1e-6
    phi = np.arctan2(nd[fl, 2], nd[fl, 0])
    d, crit = _ks_uniform((phi + np.pi) / (2 * np.pi))
    assert d < crit
    assert np.allclose(s["next_o"], s["p"].astype(np.float64) + 1e-4 * nd, atol=1e-6)


def test_emission_point_and_direction():
    sc = closed_box((0.5, 0.5, 0.5))
    ev = O.trace_events(sc, 17, N_PATHS, 1)
    assert len(ev) == N_PATHS and np.all(ev["bounce"] == 0)
    o, d = ev["o"].astype(np.float64), ev["d"].astype(np.float64)
    # origin = light point + 1e-4 * (0, -1, 0)
    assert np.allclose(o[:, 1], 0.999 - 1e-4, atol=1e-6)
    lo, w = 0.375, 0.25
    X, Z = (o[:, 0] - lo) / w, (o[:, 2] - lo) / w
    t0 = X > Z  # triangle (p0 p1 p2) of the quad; roundf(u) picks each half the time
    ok, info = _binom_ok(int(t0.sum()), N_PATHS, 0.5)
    assert ok, info
    # barycentrics in triangle 0: point = x p0 + y p1 + z p2 -> z = Z, y = X - Z
    z, y = Z[t0], X[t0] - Z[t0]
    x = 1 - y - z
    for sample in (x, y / np.maximum(1 - x, 1e-12)):
        dd, crit = _ks_uniform(np.clip(sample, 0, 1))
        assert dd < crit + 1e-4, (dd, crit)  # uniform-area sampling would give x density 2(1 - x)
    # direction density ~ cos / max|d_i|^3 on the lower hemisphere
    mu = -d[:, 1]
    assert mu.min() >= 0
    M = 2000
    MU, PH = np.meshgrid((np.arange(M) + 0.5) / M, (np.arange(2 * M) + 0.5) / (2 * M) * 2 * np.pi, indexing="ij")
    sq = np.sqrt(1 - MU ** 2)
    dens = MU / np.maximum(np.maximum(np.abs(sq * np.cos(PH)), MU), np.abs(sq * np.sin(PH))) ** 3
    edges = np.linspace(0, 1, 11)
    want = np.array([dens[(MU >= a) & (MU < b)].sum() for a, b in zip(edges[:-1], edges[1:])]) / dens.sum()
    got = np.histogram(mu, edges)[0]
    for k, p in zip(got, want):
        ok, info = _binom_ok(int(k), N_PATHS, p)
        assert ok, info
