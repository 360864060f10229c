"""CPU tests pinning the oracle's photon tracer.

The intersection core is checked against an independent float64
Moller-Trumbore brute force; the traced maps against invariants of the
reference's semantics (PhotonTracing.cu:278-375): stored power is
intensity / n_global, photons lie on diffuse surfaces (the mirror and glass
spheres never store: their Ks is 1), directions are unit length and point
into the surface, and the maps are prefixes of a longer run (path order).
"""
import numpy as np
import pytest

import pyoracle as O
from pm355.scenes import cornell_box


@pytest.fixture(scope="module")
def scene():
    return cornell_box()


def _mt_brute(scene, o, d):
    """float64 Moller-Trumbore nearest hit over all triangles."""
    tri = scene.tris9.astype(np.float64).reshape(-1, 3, 3)
    v0, v1, v2 = tri[:, 0], tri[:, 1], tri[:, 2]
    e1, e2 = v1 - v0, v2 - v0
    best_t = np.full(o.shape[0], np.inf)
    best_tri = np.full(o.shape[0], -1)
    for i in range(o.shape[0]):
        p = np.cross(d[i], e2)
        det = (e1 * p).sum(1)
        ok = np.abs(det) > 1e-12
        inv = np.where(ok, 1.0 / np.where(ok, det, 1), 0)
        s = o[i] - v0
        u = (s * p).sum(1) * inv
        q = np.cross(s, e1)
        v = (q * d[i]).sum(1) * inv
        t = (q * e2).sum(1) * inv
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= 0)
        if hit.any():
            j = np.argmin(np.where(hit, t, np.inf))
            best_t[i], best_tri[i] = t[j], j
    return best_t, best_tri


def test_philox_known_answers():
    """The generator is Philox4x32-10 as published (Salmon et al., SC'11):
    the known-answer vectors of the Random123 distribution (counter, key ->
    output) for the all-zero, all-ones and pi-digit inputs."""
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, out in kat:
        assert O.philox4x32_10(ctr, key) == out


def test_scene_hit_matches_float64_brute_force(scene):
    rng = np.random.default_rng(0)
    n = 400
    o = rng.uniform(0.05, 0.95, (n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o32, d32 = o.astype(np.float32), d.astype(np.float32)
    t, m, nrm = O.scene_hit(scene, np.concatenate([o32.T, d32.T]))
    bt, btri = _mt_brute(scene, o32.astype(np.float64), d32.astype(np.float64))
    tri_model = np.searchsorted(scene.models["tri_end"], np.arange(len(scene.tris9)), side="right")
    hit = np.isfinite(bt)
    # the box is closed except for the open front (z = 1), so some rays escape
    assert np.array_equal(m >= 0, hit)
    assert np.allclose(t[hit], bt[hit], rtol=1e-4, atol=1e-5)
    assert np.array_equal(m[hit], tri_model[btri[hit]])
    assert np.allclose(np.linalg.norm(nrm[:, hit], axis=0), 1.0, atol=1e-5)


def _on_walls_or_light(p, tol=1e-5):
    on_wall = (np.abs(p[0]) < tol) | (np.abs(p[0] - 1) < tol) | (np.abs(p[1]) < tol) | (np.abs(p[1] - 1) < tol) | \
              (np.abs(p[2]) < tol)
    on_light = (np.abs(p[1] - np.float32(0.999)) < tol) & (p[0] >= 0.375 - tol) & (p[0] <= 0.625 + tol) & \
               (p[2] >= 0.375 - tol) & (p[2] <= 0.625 + tol)
    inside = np.all((p > -tol) & (p < 1 + tol), axis=0)
    return (on_wall | on_light) & inside


def test_trace_invariants(scene):
    ng, nc = 20000, 2000
    g, c, st = O.trace_photons(scene, 1, ng, nc, 10)
    assert st[1:] == (ng, nc) and st[0] > ng
    for m in (g, c):
        assert np.all(m[3:6] == np.float32(1.0) / np.float32(ng))  # PhotonTracing.cu:369
        assert np.all(_on_walls_or_light(m[0:3]))
        ln = np.linalg.norm(m[6:9].astype(np.float64), axis=0)
        assert np.allclose(ln, 1.0, atol=1e-4)
    # photons arrive from the open side of each wall they sit on
    floor = np.abs(g[1]) < 1e-6
    assert np.all(g[7, floor] < 0)
    back = np.abs(g[2]) < 1e-6
    assert np.all(g[8, back] < 0)
    # caustics: focused by the glass sphere onto the floor / back wall region
    assert c.shape[1] == nc
    assert np.mean(np.abs(c[1]) < 1e-6) > 0.3


def test_trace_prefix_property(scene):
    g1, c1, _ = O.trace_photons(scene, 5, 4000, 300, 10)
    g2, c2, _ = O.trace_photons(scene, 5, 9000, 900, 10)
    for a, b in ((g1, g2), (c1, c2)):
        n = a.shape[1]
        assert np.array_equal(a[0:3], b[0:3, :n]) and np.array_equal(a[6:9], b[6:9, :n])
    g3, _, _ = O.trace_photons(scene, 6, 4000, 300, 10)
    assert not np.array_equal(g1[0:3], g3[0:3])


def test_trace_max_paths_and_bounces(scene):
    g, c, st = O.trace_photons(scene, 2, 100000, 100000, 10, max_paths=1000)
    assert st[0] == 1000 and st[1] < 100000 and st[2] < 1000
    g8, _, _ = O.trace_photons(scene, 2, 3000, 10, 8)
    g1, _, st1 = O.trace_photons(scene, 2, 3000, 10, 1, max_paths=20000)
    # one bounce: only first hits are stored, never caustics
    assert st1[2] == 0
