"""Synthetic, seeded inputs for the photon pipeline benchmarks and tests.

All generators are deterministic in their seed (numpy's Philox bit
generator) and return host numpy arrays in the pm355 block layouts:
photons9 = float32[9, n] (pos xyz, power rgb, inc_dir xyz), queries and
normals = float32[3, nq].
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[seed & 0xFFFFFFFFFFFFFFFF, stream]))


def _face_points(rng: np.random.Generator, n: int):
    """Uniform points on the 6 faces of [0,1]^3 and their inward normals."""
    face = rng.integers(0, 6, size=n)
    uv = rng.random((2, n), dtype=np.float32)
    axis = face // 2
    side = (face % 2).astype(np.float32)
    pos = np.empty((3, n), np.float32)
    nrm = np.zeros((3, n), np.float32)
    for a in range(3):
        sel = axis == a
        o1, o2 = [b for b in range(3) if b != a]
        pos[a, sel] = side[sel]
        pos[o1, sel] = uv[0, sel]
        pos[o2, sel] = uv[1, sel]
        nrm[a, sel] = np.where(side[sel] == 0.0, 1.0, -1.0)
    return pos, nrm


def _hemisphere_dirs(rng: np.random.Generator, nrm: np.ndarray) -> np.ndarray:
    """Directions uniform on the hemisphere about the (axis-aligned) normal."""
    n = nrm.shape[1]
    z = rng.random(n, dtype=np.float64)
    phi = rng.random(n, dtype=np.float64) * (2.0 * np.pi)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    local = np.stack([s * np.cos(phi), s * np.sin(phi), z])  # z along the normal
    out = np.empty((3, n), np.float64)
    axis = np.argmax(np.abs(nrm), axis=0)
    sign = nrm[axis, np.arange(n)]
    for a in range(3):
        sel = axis == a
        o1, o2 = [b for b in range(3) if b != a]
        out[a, sel] = local[2, sel] * sign[sel]
        out[o1, sel] = local[0, sel]
        out[o2, sel] = local[1, sel]
    return out.astype(np.float32)


def gather_microbench(seed: int = 1234, n_photons: int = 16_000_000, n_queries: int = 4_000_000,
                      n_clusters: int = 64, cluster_frac: float = 0.3, sigma: float = 0.005):
    """Photons and queries of the gather microbenchmark.

    70% of photons uniform on the 6 faces of the unit cube, 30% from
    n_clusters isotropic Gaussian clusters (sigma) on the floor y = 0 with
    centres uniform on the floor; power U[0, 1e-6] per channel; inc_dir = -w
    with w uniform on the hemisphere about the inward face normal. Queries are
    uniform on the faces with inward normals. Returns (photons9, queries3,
    normals3, aabb).
    """
    n_cl = int(round(n_photons * cluster_frac))
    n_cl -= n_cl % max(1, n_clusters)
    n_uni = n_photons - n_cl
    r_face = _rng(seed, 1)
    pos_u, nrm_u = _face_points(r_face, n_uni)
    r_cl = _rng(seed, 2)
    centres = r_cl.random((n_clusters, 2))
    per = n_cl // max(1, n_clusters)
    dev = r_cl.standard_normal((2, n_cl)) * sigma
    cid = np.repeat(np.arange(n_clusters), per)
    pos_c = np.zeros((3, n_cl), np.float32)
    pos_c[0] = np.clip(centres[cid, 0] + dev[0], 0.0, 1.0)
    pos_c[2] = np.clip(centres[cid, 1] + dev[1], 0.0, 1.0)
    nrm_c = np.zeros((3, n_cl), np.float32)
    nrm_c[1] = 1.0
    pos = np.concatenate([pos_u, pos_c], axis=1)
    nrm = np.concatenate([nrm_u, nrm_c], axis=1)
    photons = np.empty((9, n_photons), np.float32)
    photons[0:3] = pos
    photons[3:6] = (_rng(seed, 3).random((3, n_photons), dtype=np.float32) * np.float32(1e-6))
    photons[6:9] = -_hemisphere_dirs(_rng(seed, 4), nrm)
    q, qn = _face_points(_rng(seed, 5), n_queries)
    aabb = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
    return photons, q, qn, aabb


def read_photon_file(path: str) -> np.ndarray:
    """Photon map file of Initializer.cu:102-141: u64 count, then per photon
    9 float32 (pos, power, inc_dir). Returns photons9 float32[9, n]."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:8], dtype="<u8")[0])
    rec = np.frombuffer(raw[8:8 + 36 * n], dtype="<f4")
    if rec.size != 9 * n:
        raise ValueError("truncated photon file")
    return np.ascontiguousarray(rec.reshape(n, 9).T)


def write_photon_file(path: str, photons9: np.ndarray) -> None:
    """Inverse of read_photon_file (the reference has a reader only)."""
    n = photons9.shape[1]
    with open(path, "wb") as f:
        f.write(np.uint64(n).tobytes())
        f.write(np.ascontiguousarray(photons9.T, dtype="<f4").tobytes())


def batch_mix(seed: int = 7, n_jobs: int = 256):
    """Job sizes of the batched workload: per job a photon count log-uniform
    in 10k-4M, a frame from {128x128, 256x256, 640x480, 1280x720} whose pixel
    count is the job's query count, k = 64 and r uniform in [0.01, 0.05]
    (grid cell = r). Returns a list of dicts."""
    rng = _rng(seed, 11)
    res = [(128, 128), (256, 256), (640, 480), (1280, 720)]
    out = []
    for j in range(n_jobs):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(4e6))))
        w, h = res[int(rng.integers(0, 4))]
        out.append({"photons": n, "queries": w * h, "width": w, "height": h, "k": 64,
                    "radius": float(rng.uniform(0.01, 0.05)), "seed": seed * 1000 + j})
    return out


def gather_inputs_device(n_photons: int, n_queries: int, seed: int, device, n_clusters: int = 64,
                         cluster_frac: float = 0.3, sigma: float = 0.005):
    """The gather microbenchmark's distributions generated on the device with
    torch's seeded generator (for the many-job batched workload, where host
    generation would dominate): returns (photons9, queries3, normals3) as
    float32 device tensors and the unit-cube AABB."""
    import math

    import torch

    g = torch.Generator(device=device)
    g.manual_seed(int(seed))

    def faces(n):
        face = torch.randint(0, 6, (n,), generator=g, device=device)
        uv = torch.rand((2, n), generator=g, device=device)
        axis, side = face // 2, (face % 2).float()
        pos = torch.empty((3, n), device=device)
        nrm = torch.zeros((3, n), device=device)
        for a in range(3):
            sel = axis == a
            o1, o2 = [b for b in range(3) if b != a]
            pos[a] = torch.where(sel, side, pos[a])
            pos[o1] = torch.where(sel, uv[0], pos[o1])
            pos[o2] = torch.where(sel, uv[1], pos[o2])
            nrm[a] = torch.where(sel, 1.0 - 2.0 * side, nrm[a])
        return pos, nrm

    n_cl = int(round(n_photons * cluster_frac))
    n_cl -= n_cl % max(1, n_clusters)
    pos_u, nrm_u = faces(n_photons - n_cl)
    centres = torch.rand((n_clusters, 2), generator=g, device=device)
    cid = torch.arange(n_cl, device=device) // max(1, n_cl // max(1, n_clusters))
    cid = cid.clamp_max(n_clusters - 1)
    dev = torch.randn((2, n_cl), generator=g, device=device) * sigma
    pos_c = torch.zeros((3, n_cl), device=device)
    pos_c[0] = (centres[cid, 0] + dev[0]).clamp(0, 1)
    pos_c[2] = (centres[cid, 1] + dev[1]).clamp(0, 1)
    nrm_c = torch.zeros((3, n_cl), device=device)
    nrm_c[1] = 1.0
    pos = torch.cat([pos_u, pos_c], 1)
    nrm = torch.cat([nrm_u, nrm_c], 1)
    ph = torch.empty((9, n_photons), device=device)
    ph[0:3] = pos
    ph[3:6] = torch.rand((3, n_photons), generator=g, device=device) * 1e-6
    # -w, w uniform on the hemisphere about the axis-aligned normal
    z = torch.rand(n_photons, generator=g, device=device)
    phi = torch.rand(n_photons, generator=g, device=device) * (2 * math.pi)
    s = (1 - z * z).clamp_min(0).sqrt()
    loc = torch.stack([s * phi.cos(), s * phi.sin(), z])
    axis = nrm.abs().argmax(0)
    sign = nrm.gather(0, axis[None])[0]
    w = torch.empty((3, n_photons), device=device)
    for a in range(3):
        sel = axis == a
        o1, o2 = [b for b in range(3) if b != a]
        w[a] = torch.where(sel, loc[2] * sign, w[a])
        w[o1] = torch.where(sel, loc[0], w[o1])
        w[o2] = torch.where(sel, loc[1], w[o2])
    ph[6:9] = -w
    q, qn = faces(n_queries)
    return ph.contiguous(), q.contiguous(), qn.contiguous(), (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
