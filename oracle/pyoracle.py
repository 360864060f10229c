"""ctypes wrapper over liboracle.so — TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's semantics. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker / CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_U32 = C.c_uint32
_I32 = C.c_int32
_F32 = C.c_float
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"oracle not built at {LIB_PATH} (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_scan_u32.argtypes = [_P, _U32, _I32, _P, _P]
        L.orc_grid_plan.argtypes = [_P, _U32, _P, _F32, _I32, _I32, _P, _P, _P]
        L.orc_grid_plan.restype = _I32
        L.orc_grid_fill.argtypes = [_P, _U32, _P, _F32, _P, _I32, _P, _P, _P, _P]
        L.orc_grid_fill.restype = _I32
        L.orc_grid_knn.argtypes = [_P, _P, _P, _U32, _P, _F32, _P, _I32, _P, _U32, _F32, _I32, _I32,
                                   _P, _P, _P, _P, _P, _P]
        L.orc_grid_knn.restype = _I32
        L.orc_knn_brute.argtypes = [_P, _U32, _P, _U32, _F32, _I32, _P, _P, _P]
        L.orc_knn_brute.restype = _I32
        L.orc_radiance.argtypes = [_P, _P, _P, _U32, _I32, _P, _P, _U32, _P, _P, _P]
        L.orc_radiance.restype = _I32
        L.orc_kdtree_build.argtypes = [_P, _U32, _U32, _P, _P]
        L.orc_kdtree_build.restype = _I32
        L.orc_kdtree_knn.argtypes = [_P, _U32, _P, _P, _U32, _P, _U32, _F32, _I32, _I32, _P, _P, _P]
        L.orc_kdtree_knn.restype = _I32
        L.orc_model_aabb.argtypes = [_P, _P, _I32, _P]
        L.orc_model_aabb.restype = _I32
        L.orc_radius_threshold.argtypes = [_F32]
        L.orc_radius_threshold.restype = _F32
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def scan_u32(x, inclusive=False):
    x = _c(x, np.uint32)
    out = np.empty_like(x)
    tot = np.zeros(1, np.uint64)
    lib().orc_scan_u32(x.ctypes.data, x.size, int(inclusive), out.ctypes.data, tot.ctypes.data)
    return out, int(tot[0])


class OracleGrid:
    def __init__(self, photons9, aabb, cell_size=0.0, desired=1000, layout=0):
        ph = _c(photons9, np.float32)
        self.n = ph.shape[1]
        ab = _c(aabb, np.float32)
        cell = np.zeros(1, np.float32)
        dims = np.zeros(3, np.int32)
        nc = np.zeros(1, np.uint32)
        rc = lib().orc_grid_plan(ph.ctypes.data, self.n, ab.ctypes.data, float(cell_size), int(desired),
                                 int(layout), cell.ctypes.data, dims.ctypes.data, nc.ctypes.data)
        if rc != 0:
            raise ValueError(f"orc_grid_plan rc={rc}")
        self.aabb, self.cell, self.dims, self.n_cells, self.layout = ab, float(cell[0]), dims, int(nc[0]), layout
        self.counts = np.empty(self.n_cells, np.uint32)
        self.starts = np.empty(self.n_cells, np.uint32)
        self.photons = np.empty((9, self.n), np.float32)
        self.perm = np.empty(self.n, np.uint32)
        rc = lib().orc_grid_fill(ph.ctypes.data, self.n, ab.ctypes.data, self.cell, dims.ctypes.data, int(layout),
                                 self.counts.ctypes.data, self.starts.ctypes.data, self.photons.ctypes.data,
                                 self.perm.ctypes.data)
        if rc != 0:
            raise ValueError(f"orc_grid_fill rc={rc}")

    def knn(self, queries3, r, k, threads=0):
        q = _c(queries3, np.float32)
        nq = q.shape[1]
        idx = np.empty((nq, k), np.uint32)
        dist = np.empty((nq, k), np.float32)
        cnt = np.empty(nq, np.uint32)
        st = np.zeros(3, np.uint64)
        rc = lib().orc_grid_knn(self.counts.ctypes.data, self.starts.ctypes.data, self.photons.ctypes.data, self.n,
                                self.aabb.ctypes.data, self.cell, self.dims.ctypes.data, self.layout, q.ctypes.data,
                                nq, float(r), int(k), int(threads), idx.ctypes.data, dist.ctypes.data,
                                cnt.ctypes.data, st[0:].ctypes.data, st[1:].ctypes.data, st[2:].ctypes.data)
        if rc != 0:
            raise ValueError(f"orc_grid_knn rc={rc}")
        self.last_stats = dict(cells=int(st[0]), cands=int(st[1]), found=int(st[2]))
        return idx, dist, cnt


def knn_brute(pos3, queries3, r, k):
    p = _c(pos3, np.float32)
    q = _c(queries3, np.float32)
    nq = q.shape[1]
    idx = np.empty((nq, k), np.uint32)
    dist = np.empty((nq, k), np.float32)
    cnt = np.empty(nq, np.uint32)
    rc = lib().orc_knn_brute(p.ctypes.data, p.shape[1], q.ctypes.data, nq, float(r), int(k), idx.ctypes.data,
                             dist.ctypes.data, cnt.ctypes.data)
    assert rc == 0
    return idx, dist, cnt


def radiance(idx, dist, cnt, photons9, normals3, k):
    idx = _c(idx, np.uint32)
    dist = _c(dist, np.float32)
    cnt = _c(cnt, np.uint32)
    ph = _c(photons9, np.float32)
    n = ph.shape[1]
    nrm = _c(normals3, np.float32)
    nq = nrm.shape[1]
    out32 = np.empty((3, nq), np.float32)
    out64 = np.empty((3, nq), np.float64)
    power = np.ascontiguousarray(ph[3:6])
    dirs = np.ascontiguousarray(ph[6:9])
    rc = lib().orc_radiance(idx.ctypes.data, dist.ctypes.data, cnt.ctypes.data, nq, int(k),
                            power.ctypes.data, dirs.ctypes.data, n, nrm.ctypes.data,
                            out32.ctypes.data, out64.ctypes.data)
    assert rc == 0
    return out32, out64


def kdtree_build(pos3):
    p = _c(pos3, np.float32)
    n = p.shape[1]
    L = int(np.ceil(np.log2(n + 1.0))) if n > 0 else 0
    cap = (1 << L) - 1 if n > 0 else 0
    node_idx = np.empty(max(cap, 1), np.uint32)
    node_plane = np.empty(max(cap, 1), np.int8)
    rc = lib().orc_kdtree_build(p.ctypes.data, n, cap, node_idx.ctypes.data, node_plane.ctypes.data)
    assert rc == 0, rc
    return node_idx[:cap], node_plane[:cap]


def kdtree_knn(pos3, node_idx, node_plane, queries3, r, k, threads=0):
    p = _c(pos3, np.float32)
    q = _c(queries3, np.float32)
    ni = _c(node_idx, np.uint32)
    npl = _c(node_plane, np.int8)
    nq = q.shape[1]
    idx = np.empty((nq, k), np.uint32)
    dist = np.empty((nq, k), np.float32)
    cnt = np.empty(nq, np.uint32)
    rc = lib().orc_kdtree_knn(p.ctypes.data, p.shape[1], ni.ctypes.data, npl.ctypes.data, ni.size, q.ctypes.data,
                              nq, float(r), int(k), int(threads), idx.ctypes.data, dist.ctypes.data, cnt.ctypes.data)
    assert rc == 0
    return idx, dist, cnt


def model_aabb(verts9, tri_begin):
    v = _c(verts9, np.float32)
    tb = _c(tri_begin, np.uint32)
    nm = tb.size - 1
    out = np.empty((nm, 6), np.float32)
    lib().orc_model_aabb(v.ctypes.data, tb.ctypes.data, nm, out.ctypes.data)
    return out


def radius_threshold(r):
    return float(np.float32(lib().orc_radius_threshold(float(r))))


def set_threads(n):
    """Threads of the oracle's OpenMP loops that take no thread argument
    (the tracer); n <= 0 restores the default."""
    L = lib()
    L.orc_set_threads.argtypes = [_I32]
    L.orc_set_threads.restype = _I32
    return int(L.orc_set_threads(int(n)))


def philox4x32_10(ctr, key):
    """One Philox4x32-10 block of the tracer's generator: 4 u32 counter words,
    2 u32 key words -> 4 u32."""
    L = lib()
    L.orc_philox4x32_10.argtypes = [_P, _P, _P]
    L.orc_philox4x32_10.restype = None
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    L.orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return tuple(int(v) for v in out)


def trace_photons(scene, seed, n_global, n_caustic, max_bounces=10, max_paths=1 << 40):
    """Oracle photon tracing on a pm355.scenes.Scene; returns (global9,
    caustic9, stats) with stats = (paths, n_global_stored, n_caustic_stored)."""
    L = lib()
    if not getattr(L, "_trace_sig", False):
        L.orc_trace_photons.argtypes = [_P, _P, _U32, _P, _I32, _I32, _P, _F32, C.c_uint64, _U32, _U32, _I32,
                                        C.c_uint64, _P, _P, _P]
        L.orc_trace_photons.restype = _I32
        L._trace_sig = True
    tris = _c(scene.tris9, np.float32)
    nrms = _c(scene.normals9, np.float32)
    models = np.ascontiguousarray(scene.models)
    inten = _c(scene.light_intensity, np.float32)
    g = np.zeros((9, max(n_global, 1)), np.float32)
    c = np.zeros((9, max(n_caustic, 1)), np.float32)
    st = np.zeros(3, np.uint64)
    rc = L.orc_trace_photons(tris.ctypes.data, nrms.ctypes.data, tris.shape[0], models.ctypes.data, len(models),
                             int(scene.light_model), inten.ctypes.data, float(scene.default_refr_index), int(seed),
                             int(n_global), int(n_caustic), int(max_bounces), int(max_paths), g.ctypes.data,
                             c.ctypes.data, st.ctypes.data)
    if rc != 0:
        raise ValueError(f"orc_trace_photons rc={rc}")
    return g[:, :n_global], c[:, :n_caustic], tuple(int(v) for v in st)


def scene_hit(scene, rays6):
    """Nearest hit of each ray (6 planes: origin xyz, direction xyz):
    (t, model, normal3); t = -1 and model = -1 on a miss."""
    L = lib()
    L.orc_scene_hit.argtypes = [_P, _P, _U32, _P, _I32, _P, _U32, _P, _P, _P]
    L.orc_scene_hit.restype = _I32
    tris = _c(scene.tris9, np.float32)
    nrms = _c(scene.normals9, np.float32)
    models = np.ascontiguousarray(scene.models)
    r = _c(rays6, np.float32)
    nr = r.shape[1]
    t = np.empty(nr, np.float32)
    m = np.empty(nr, np.int32)
    n = np.empty((3, nr), np.float32)
    rc = L.orc_scene_hit(tris.ctypes.data, nrms.ctypes.data, tris.shape[0], models.ctypes.data, len(models),
                         r.ctypes.data, nr, t.ctypes.data, m.ctypes.data, n.ctypes.data)
    assert rc == 0
    return t, m, n


def render(scene, og, oc, camera, max_depth=4, k_global=50, r_global=0.05, k_caustic=50, r_caustic=0.05,
           flags=0, threads=0):
    """Oracle frame (render_cpu semantics) on OracleGrid og (global) and oc
    (caustic, or None). Returns (rgb3 float32[3, H*W], rgb8 uint8[H*W*3], hits)."""
    L = lib()
    L.orc_render.argtypes = [_P, _P, _U32, _P, _I32, _F32, _P, _P, _P, _U32, _P, _F32, _P, _I32,
                             _P, _P, _P, _U32, _P, _F32, _P, _I32, _P, _P, _P, _I32, _P, _P, _P]
    L.orc_render.restype = _I32
    tris = _c(scene.tris9, np.float32)
    nrms = _c(scene.normals9, np.float32)
    models = np.ascontiguousarray(scene.models)
    cam = np.ascontiguousarray(camera)
    W, H = int(cam["width"].item()), int(cam["height"].item())
    ip = np.array([max_depth, k_global, k_caustic, flags], np.int32)
    fp = np.array([r_global, r_caustic], np.float32)
    rgb = np.empty((3, W * H), np.float32)
    rgb8 = np.empty(W * H * 3, np.uint8)
    st = np.zeros(1, np.uint64)
    z = np.zeros(8, np.uint32)
    if oc is None:
        c_args = (z.ctypes.data, z.ctypes.data, z.ctypes.data, 0, og.aabb.ctypes.data, 1.0, og.dims.ctypes.data, 0)
    else:
        c_args = (oc.counts.ctypes.data, oc.starts.ctypes.data, oc.photons.ctypes.data, oc.n, oc.aabb.ctypes.data,
                  oc.cell, oc.dims.ctypes.data, oc.layout)
    if oc is None:
        c_args = (None, None, None, 0) + c_args[4:]
    rc = L.orc_render(tris.ctypes.data, nrms.ctypes.data, tris.shape[0], models.ctypes.data, len(models),
                      float(scene.default_refr_index), og.counts.ctypes.data, og.starts.ctypes.data,
                      og.photons.ctypes.data, og.n, og.aabb.ctypes.data, og.cell, og.dims.ctypes.data, og.layout,
                      *c_args, cam.ctypes.data, ip.ctypes.data, fp.ctypes.data, int(threads), rgb.ctypes.data,
                      rgb8.ctypes.data, st.ctypes.data)
    if rc != 0:
        raise ValueError(f"orc_render rc={rc}")
    return rgb, rgb8, int(st[0])


TRACE_EVENT_DTYPE = np.dtype([
    ("path", "<u4"), ("bounce", "<i4"), ("model", "<i4"), ("dest", "<i4"), ("stored", "<i4"), ("flags", "<i4"),
    ("o", "<f4", 3), ("d", "<f4", 3), ("p", "<f4", 3), ("n", "<f4", 3), ("next_o", "<f4", 3), ("next_d", "<f4", 3),
    ("n1", "<f4"), ("n2", "<f4"),
])
# event kinds (the Path values of PathOperator.cuh)
EV_DIFFUSE, EV_SPECULAR, EV_ABSORB, EV_REFRACT = 0, 1, 2, 3


def trace_events(scene, seed, n_paths, max_bounces=10, n_global=1, cap=None):
    """Every bounce event of paths [0, n_paths) as the oracle tracer decides
    them (TRACE_EVENT_DTYPE records in (path, bounce) order)."""
    L = lib()
    L.orc_trace_events.argtypes = [_P, _P, _U32, _P, _I32, _I32, _P, _F32, C.c_uint64, _U32, _I32, C.c_uint64,
                                   _P, C.c_uint64, _P]
    L.orc_trace_events.restype = _I32
    tris = _c(scene.tris9, np.float32)
    nrms = _c(scene.normals9, np.float32)
    models = np.ascontiguousarray(scene.models)
    inten = _c(scene.light_intensity, np.float32)
    cap = int(cap or n_paths * (max_bounces + 1))
    out = np.empty(cap, TRACE_EVENT_DTYPE)
    n = np.zeros(1, np.uint64)
    rc = L.orc_trace_events(tris.ctypes.data, nrms.ctypes.data, tris.shape[0], models.ctypes.data, len(models),
                            int(scene.light_model), inten.ctypes.data, float(scene.default_refr_index), int(seed),
                            int(n_global), int(max_bounces), int(n_paths), out.ctypes.data, cap, n.ctypes.data)
    if rc != 0:
        raise ValueError(f"orc_trace_events rc={rc} (events {int(n[0])}, cap {cap})")
    return out[:int(n[0])]
