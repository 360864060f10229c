"""Python harness over the pm355 C ABI (include/pm355.h).

This module is plumbing for tests and benchmarks: it loads the in-tree
``libpm355.so`` with ctypes and passes torch device tensors' pointers to it.
Every computation happens in the HIP library; there is no Python or CPU
fallback, and loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG_DIR), "libpm355.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_PKG_DIR)), "include", "pm355.h")

PM_OK = 0
PM_ERR_INVALID = -1
PM_ERR_HIP = -2
PM_ERR_NOMEM = -3
PM_ERR_CAPACITY = -4
PM_ERR_UNSUPPORTED = -5
PM_ERR_NO_DEVICE = -6

LAYOUT_LINEAR = 0
LAYOUT_MORTON = 1
KNN_SORT_QUERIES = 1
KNN_NO_XCD_REMAP = 2
KNN_TWO_PASS = 1 << 4
KNN_LARGE_BUFFER = 1 << 5
KNN_BATCH_QUEUE = 1 << 6
KNN_XCD_CONTIGUOUS = 1 << 7
KNN_INNER_BOX = 1 << 8
KNN_TUNED = 1 << 9
BOX_REFERENCE = 1
# gather variants (tuning flags): the default is the 512-key buffer with
# XCD-interleaved chunks of workgroups
KNN_VARIANTS = {"stream512_chunked": 0, "stream512": KNN_XCD_CONTIGUOUS, "stream1024": KNN_LARGE_BUFFER,
                "two_pass": KNN_TWO_PASS, "inner": KNN_INNER_BOX}

_P = C.c_void_p
_U32 = C.c_uint32
_I32 = C.c_int32
_F32 = C.c_float


class PMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pm355 error {code}: {msg}")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [
        ("aabb", C.c_float * 6),
        ("cell_size", C.c_float),
        ("dims", C.c_int32 * 3),
        ("layout", C.c_int32),
        ("n_cells", C.c_uint32),
        ("n_photons", C.c_uint32),
        ("d_cell_starts", C.c_void_p),
        ("d_cell_sizes", C.c_void_p),
        ("d_photons", C.c_void_p),
        ("d_perm", C.c_void_p),
    ]


_SIGS = {
    "pm_version": (C.c_char_p, []),
    "pm_ctx_create": (_I32, [_I32, _P, C.POINTER(_P)]),
    "pm_ctx_destroy": (_I32, [_P]),
    "pm_ctx_set_stream": (_I32, [_P, _P]),
    "pm_ctx_sync": (_I32, [_P]),
    "pm_ctx_trim": (_I32, [_P]),
    "pm_last_error": (C.c_char_p, [_P]),
    "pm_scan_u32": (_I32, [_P, _P, _P, _U32, _I32, _P]),
    "pm_grid_build": (_I32, [_P, _P, _U32, C.POINTER(_F32), _F32, _I32, _I32, C.POINTER(_P)]),
    "pm_grid_rebuild": (_I32, [_P, _P, _P, _U32]),
    "pm_grid_describe": (_I32, [_P, C.POINTER(GridDesc)]),
    "pm_grid_copy_to_host": (_I32, [_P, _P, _P, _P, _P, _P]),
    "pm_grid_destroy": (_I32, [_P]),
    "pm_grid_knn": (_I32, [_P, _P, _P, _U32, _F32, _I32, _U32, _P, _P, _P]),
    "pm_radiance": (_I32, [_P, _P, _P, _P, _P, _P, _U32, _I32, _P]),
    "pm_grid_gather_radiance": (_I32, [_P, _P, _P, _P, _U32, _F32, _I32, _U32, _P, _P]),
    "pm_grid_gather_radiance_noalloc": (_I32, [_P, _P, _P, _P, _U32, _F32, _I32, _U32, _P, _P]),
    "pm_grid_query_order": (_I32, [_P, _P, _P, _U32, _P]),
    "pm_grid_box_stats": (_I32, [_P, _P, _P, _U32, _F32, _U32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "pm_kdtree_build": (_I32, [_P, _P, _U32, C.POINTER(_P)]),
    "pm_kdtree_describe": (_I32, [_P, _P]),
    "pm_kdtree_copy_to_host": (_I32, [_P, _P, _P, _P]),
    "pm_kdtree_destroy": (_I32, [_P]),
    "pm_kdtree_knn": (_I32, [_P, _P, _P, _U32, _F32, _I32, _P, _P, _P]),
    "pm_accel_build": (_I32, [_P, _P, C.POINTER(_P)]),
    "pm_gather_batch": (_I32, [_P, _P, _U32, _U32]),
    "pm_frame_create": (_I32, [_P, C.POINTER(_P)]),
    "pm_frame_destroy": (_I32, [_P]),
    "pm_frame_rays": (_I32, [_P, _P, _P, _P, _P, C.POINTER(_U32)]),
    "pm_frame_gather_jobs": (_I32, [_P, _P, _P, _P, C.POINTER(_U32)]),
    "pm_frame_finish": (_I32, [_P, _P, _P, _P, _P]),
    "pm_server_create": (_I32, [_P, C.c_size_t, C.POINTER(_P)]),
    "pm_server_destroy": (_I32, [_P]),
    "pm_server_submit": (_I32, [_P, _P, C.POINTER(C.c_uint64)]),
    "pm_server_step": (_I32, [_P, C.POINTER(_U32)]),
    "pm_server_set_workers": (_I32, [_P, _I32]),
    "pm_tune_gather": (_I32, [_P, _P, _U32, _P, _P, _P, _U32, _F32, _I32, _P]),
    "pm_tuning_lookup": (_I32, [_P, _U32, _P, _F32, _I32, _P]),
    "pm_tuning_table": (_I32, [_P, _P, _U32]),
    "pm_frame_copy_queries": (_I32, [_P, _P, _P, _P, _U32, C.POINTER(_U32)]),
    "pm_server_stats": (_I32, [_P, _P]),
    "pm_server_latencies": (_I32, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "pm_accel_destroy": (_I32, [_P]),
    "pm_model_aabb": (_I32, [_P, _P, _P, _I32, _P]),
    "pm_render_frame": (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pm_trace_photons": (_I32, [_P, _P, C.c_uint64, _U32, _U32, _I32, C.c_uint64, _P, _P, _P]),
    "pm_pipeline_create": (_I32, [_P, _P, C.POINTER(_P)]),
    "pm_pipeline_frame": (_I32, [_P, C.c_uint64, _P, _P]),
    "pm_pipeline_status": (_I32, [_P, _P]),
    "pm_pipeline_grids": (_I32, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "pm_pipeline_destroy": (_I32, [_P]),
    "pm_photon_file_count": (_I32, [_P, C.c_char_p, C.POINTER(C.c_uint64)]),
    "pm_photon_file_read": (_I32, [_P, C.c_char_p, _P, _U32, C.POINTER(_U32)]),
    "pm_photon_file_write": (_I32, [_P, C.c_char_p, _P, _U32]),
}


class ModelRec(C.Structure):
    _fields_ = [("tri_begin", C.c_uint32), ("tri_end", C.c_uint32), ("kd", C.c_float * 3), ("ks", C.c_float * 3),
                ("ke", C.c_float * 3), ("refr_index", C.c_float), ("opaque", C.c_float), ("shininess", C.c_float),
                ("id", C.c_int32), ("pad", C.c_int32), ("aabb", C.c_float * 6)]


class SceneRec(C.Structure):
    _fields_ = [("d_tris9", C.c_void_p), ("d_normals9", C.c_void_p), ("n_tris", C.c_uint32),
                ("h_models", C.c_void_p), ("n_models", C.c_int32), ("light_model", C.c_int32),
                ("light_intensity", C.c_float * 3), ("default_refr_index", C.c_float), ("accel", C.c_void_p)]


class GatherJob(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("d_queries3", C.c_void_p), ("nq", C.c_uint32), ("radius", C.c_float),
                ("k", C.c_int32), ("d_idx", C.c_void_p), ("d_dist", C.c_void_p), ("d_count", C.c_void_p),
                ("d_normals3", C.c_void_p), ("d_rgb3", C.c_void_p), ("flags", C.c_uint32)]


class GatherTuning(C.Structure):
    _fields_ = [("cell_over_r", C.c_float), ("flags", C.c_uint32), ("ms", C.c_float), ("ms_default", C.c_float),
                ("found", C.c_int32)]


class CameraRec(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("viewport_width", C.c_float), ("viewport_height", C.c_float),
                ("focal_length", C.c_float), ("width", C.c_int32), ("height", C.c_int32), ("pad", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("aabb", C.c_float * 6), ("cell_size", C.c_float), ("n_global", C.c_uint32),
                ("n_caustic", C.c_uint32), ("max_bounces", C.c_int32), ("max_paths", C.c_uint64),
                ("seed", C.c_uint64), ("camera", CameraRec), ("params", C.c_int32 * 7), ("frames", C.c_uint32),
                ("d_rgb8", C.c_void_p)]


class PipelineDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("aabb", C.c_float * 6), ("cell_global", C.c_float),
                ("cell_caustic", C.c_float), ("n_global", C.c_uint32), ("n_caustic", C.c_uint32),
                ("max_bounces", C.c_int32), ("max_paths", C.c_uint64), ("plan_seed", C.c_uint64),
                ("camera", CameraRec), ("params", C.c_int32 * 7)]


class PipelineInfo(C.Structure):
    _fields_ = [("scheduled_paths", C.c_uint64), ("n_global", C.c_uint32), ("n_caustic", C.c_uint32),
                ("complete", C.c_uint32), ("hits", C.c_uint64)]


class ServerInfo(C.Structure):
    _fields_ = [("queued", C.c_uint32), ("active", C.c_uint32), ("finished", C.c_uint32), ("frames", C.c_uint64),
                ("queries", C.c_uint64), ("bytes_in_use", C.c_uint64)]


class TraceStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("n_global", C.c_uint32), ("n_caustic", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("k_global", C.c_int32), ("r_global", C.c_float),
                ("k_caustic", C.c_int32), ("r_caustic", C.c_float), ("flags", C.c_uint32),
                ("gather_flags", C.c_uint32)]


class FrameStats(C.Structure):
    _fields_ = [("hits", C.c_uint64), ("pixels", C.c_uint32)]


RENDER_REFERENCE_CAUSTIC = 1


class KdTreeDesc(C.Structure):
    _fields_ = [
        ("n_photons", C.c_uint32),
        ("n_slots", C.c_uint32),
        ("levels", C.c_uint32),
        ("d_node_idx", C.c_void_p),
        ("d_node_plane", C.c_void_p),
        ("d_photons", C.c_void_p),
    ]

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libpm355.so; raises if it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libpm355.so not built at {path}; run __graft_entry__.build()")
    # torch first: its HIP runtime must be the process's one before the
    # library binds to a HIP runtime of its own
    import torch  # noqa: F401

    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols(header: str = HEADER_PATH) -> list[str]:
    """Names of the functions declared in include/pm355.h."""
    import re

    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)))


def _ptr(t) -> int:
    if t is None:
        return None
    return t.data_ptr()


def _f32_dev(a, dev):
    import torch

    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


class Context:
    """A pm_ctx bound to a torch device and that device's current stream."""

    def __init__(self, device: int = 0):
        import torch

        self.lib = load_library()
        self.device = torch.device("cuda", device)
        self.stream = torch.cuda.current_stream(self.device)
        h = _P()
        rc = self.lib.pm_ctx_create(device, _P(self.stream.cuda_stream), C.byref(h))
        if rc != PM_OK:
            raise PMError(rc, "pm_ctx_create failed (no gfx950 device?)")
        self.h = h

    def check(self, rc: int):
        if rc != PM_OK:
            raise PMError(rc, self.lib.pm_last_error(self.h).decode())

    def sync(self):
        self.check(self.lib.pm_ctx_sync(self.h))

    def trim(self):
        """Free the device blocks that destroyed grids and frames left in
        this context's cache (pm_ctx_trim)."""
        self.check(self.lib.pm_ctx_trim(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.pm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scan ---------------------------------------------------------------
    def scan_u32(self, x, inclusive: bool = False, with_total: bool = False):
        import torch

        x = x.to(self.device).contiguous()
        out = torch.empty_like(x)
        tot = torch.zeros(1, dtype=torch.int64, device=self.device) if with_total else None
        self.check(self.lib.pm_scan_u32(self.h, _ptr(x), _ptr(out), x.numel(), int(inclusive), _ptr(tot)))
        return (out, tot) if with_total else out

    # -- grid -----------------------------------------------------------------
    def grid_build(self, photons9, aabb, cell_size: float = 0.0, desired: int = 1000,
                   layout: int = LAYOUT_LINEAR) -> "Grid":
        import torch

        ph = _f32_dev(photons9, self.device)
        assert ph.dim() == 2 and ph.shape[0] == 9, "photons9 must be [9, n]"
        n = ph.shape[1]
        ab = (C.c_float * 6)(*[float(v) for v in aabb])
        g = _P()
        self.check(self.lib.pm_grid_build(self.h, _ptr(ph), n, ab, float(cell_size), int(desired), int(layout),
                                          C.byref(g)))
        return Grid(self, g, keep=ph)

    def knn(self, grid: "Grid", queries3, radius: float, k: int, flags: int = 0):
        import torch

        q = _f32_dev(queries3, self.device)
        nq = q.shape[1]
        idx = torch.empty((nq, k), dtype=torch.int32, device=self.device)
        dist = torch.empty((nq, k), dtype=torch.float32, device=self.device)
        cnt = torch.empty((nq,), dtype=torch.int32, device=self.device)
        self.check(self.lib.pm_grid_knn(self.h, grid.h, _ptr(q), nq, float(radius), int(k), int(flags),
                                        _ptr(idx), _ptr(dist), _ptr(cnt)))
        return idx, dist, cnt

    def radiance(self, grid: "Grid", idx, dist, cnt, normals3, k: int):
        import torch

        nrm = _f32_dev(normals3, self.device)
        nq = nrm.shape[1]
        rgb = torch.empty((3, nq), dtype=torch.float32, device=self.device)
        self.check(self.lib.pm_radiance(self.h, grid.h, _ptr(idx), _ptr(dist), _ptr(cnt), _ptr(nrm), nq, int(k),
                                        _ptr(rgb)))
        return rgb

    def gather_radiance(self, grid: "Grid", queries3, normals3, radius: float, k: int, flags: int = 0,
                        want_count: bool = False):
        import torch

        q = _f32_dev(queries3, self.device)
        nrm = _f32_dev(normals3, self.device)
        nq = q.shape[1]
        rgb = torch.empty((3, nq), dtype=torch.float32, device=self.device)
        cnt = torch.empty((nq,), dtype=torch.int32, device=self.device) if want_count else None
        self.check(self.lib.pm_grid_gather_radiance(self.h, grid.h, _ptr(q), _ptr(nrm), nq, float(radius), int(k),
                                                    int(flags), _ptr(rgb), _ptr(cnt)))
        return (rgb, cnt) if want_count else rgb

    def gather_radiance_noalloc(self, grid: "Grid", queries3, normals3, radius: float, k: int, flags: int = 0,
                                rgb=None, count=None):
        """pm_grid_gather_radiance_noalloc into caller-provided outputs
        (rgb: (3, nq) float32, count: (nq,) int32 or None) on device."""
        q = _f32_dev(queries3, self.device)
        nrm = _f32_dev(normals3, self.device)
        nq = q.shape[1]
        self.check(self.lib.pm_grid_gather_radiance_noalloc(self.h, grid.h, _ptr(q), _ptr(nrm), nq, float(radius),
                                                            int(k), int(flags), _ptr(rgb), _ptr(count)))
        return rgb, count

    def tune_gather(self, photons9, aabb, queries3, normals3, radius: float, k: int) -> dict:
        """pm_tune_gather: time the cell sizes and gather variants on these
        photons and queries; the choice is remembered for PM_KNN_TUNED."""
        ph = _f32_dev(photons9, self.device)
        q = _f32_dev(queries3, self.device)
        nrm = _f32_dev(normals3, self.device)
        ab = (C.c_float * 6)(*[float(v) for v in aabb])
        t = GatherTuning()
        self.check(self.lib.pm_tune_gather(self.h, _ptr(ph), ph.shape[1], ab, _ptr(q), _ptr(nrm), q.shape[1],
                                           float(radius), int(k), C.byref(t)))
        return {f: getattr(t, f) for f, _ in GatherTuning._fields_}

    TUNE_FRACS = (0.25, 1.0 / 3.0, 0.4, 0.5, 2.0 / 3.0, 1.0)
    TUNE_VARIANTS = ("stream512_chunked", "stream512", "stream1024", "two_pass", "inner")

    def tuning_table(self) -> dict:
        """The last tune_gather's measurements: {cell/r: {variant: ms}}
        (grid rebuild + gather; configurations not measured are left out)."""
        buf = (C.c_float * 30)()
        self.check(self.lib.pm_tuning_table(self.h, buf, 30))
        out = {}
        for fi, f in enumerate(self.TUNE_FRACS):
            row = {v: buf[fi * 5 + vi] for vi, v in enumerate(self.TUNE_VARIANTS) if buf[fi * 5 + vi] >= 0}
            if row:
                out[f"{f:.3f}"] = row
        return out

    def tuning_lookup(self, n: int, aabb, radius: float, k: int) -> dict:
        ab = (C.c_float * 6)(*[float(v) for v in aabb])
        t = GatherTuning()
        self.check(self.lib.pm_tuning_lookup(self.h, int(n), ab, float(radius), int(k), C.byref(t)))
        return {f: getattr(t, f) for f, _ in GatherTuning._fields_}

    def gather_batch(self, jobs, flags: int = 0, want_knn: bool = False, want_count: bool = True):
        """Batched gather over many (grid, queries3, normals3, radius, k) jobs in
        one persistent launch. Returns per job a dict with "rgb" (when normals
        are given), "count", and "idx"/"dist" when want_knn."""
        import torch

        recs = (GatherJob * max(len(jobs), 1))()
        outs, keep = [], []
        for i, job in enumerate(jobs):
            # (grid, queries3, normals3, radius, k[, flags of this job])
            grid, q3, n3, radius, k = job[:5]
            q = _f32_dev(q3, self.device)
            nq = q.shape[1]
            o = {}
            r = recs[i]
            r.grid, r.d_queries3, r.nq, r.radius, r.k = grid.h, _ptr(q), nq, float(radius), int(k)
            r.flags = int(job[5]) if len(job) > 5 else 0
            keep.append(q)
            if n3 is not None:
                nrm = _f32_dev(n3, self.device)
                keep.append(nrm)
                o["rgb"] = torch.empty((3, nq), dtype=torch.float32, device=self.device)
                r.d_normals3, r.d_rgb3 = _ptr(nrm), _ptr(o["rgb"])
            if want_count or n3 is None:
                o["count"] = torch.empty((nq,), dtype=torch.int32, device=self.device)
                r.d_count = _ptr(o["count"])
            if want_knn:
                o["idx"] = torch.empty((nq, k), dtype=torch.int32, device=self.device)
                o["dist"] = torch.empty((nq, k), dtype=torch.float32, device=self.device)
                r.d_idx, r.d_dist = _ptr(o["idx"]), _ptr(o["dist"])
            outs.append(o)
        self.check(self.lib.pm_gather_batch(self.h, C.cast(recs, C.c_void_p), len(jobs), int(flags)))
        return outs

    # -- kd-tree / model AABB ---------------------------------------------------
    def kdtree_build(self, photons9) -> "KdTree":
        ph = _f32_dev(photons9, self.device)
        assert ph.dim() == 2 and ph.shape[0] == 9
        t = _P()
        self.check(self.lib.pm_kdtree_build(self.h, _ptr(ph), ph.shape[1], C.byref(t)))
        return KdTree(self, t, keep=ph)

    def kdtree_knn(self, tree: "KdTree", queries3, radius: float, k: int):
        import torch

        q = _f32_dev(queries3, self.device)
        nq = q.shape[1]
        idx = torch.empty((nq, k), dtype=torch.int32, device=self.device)
        dist = torch.empty((nq, k), dtype=torch.float32, device=self.device)
        cnt = torch.empty((nq,), dtype=torch.int32, device=self.device)
        self.check(self.lib.pm_kdtree_knn(self.h, tree.h, _ptr(q), nq, float(radius), int(k), _ptr(idx), _ptr(dist),
                                          _ptr(cnt)))
        return idx, dist, cnt

    def model_aabb(self, tris9, tri_begin):
        import torch

        tr = _f32_dev(tris9, self.device)
        tb = torch.from_numpy(np.ascontiguousarray(tri_begin, dtype=np.uint32).view(np.int32)).to(self.device)
        nm = tb.numel() - 1
        out = torch.empty((nm, 6), dtype=torch.float32, device=self.device)
        self.check(self.lib.pm_model_aabb(self.h, _ptr(tr), _ptr(tb), nm, _ptr(out)))
        return out

    def query_order(self, grid: "Grid", queries3):
        import torch

        q = _f32_dev(queries3, self.device)
        order = torch.empty((q.shape[1],), dtype=torch.int32, device=self.device)
        self.check(self.lib.pm_grid_query_order(self.h, grid.h, _ptr(q), q.shape[1], _ptr(order)))
        return order

    # -- photon-map files (Initializer.cu:102-141) --------------------------------
    def photon_file_read(self, path: str):
        """A photon-map file into a new device photons9 tensor [9, n]."""
        import torch

        n = C.c_uint64()
        self.check(self.lib.pm_photon_file_count(self.h, os.fsencode(path), C.byref(n)))
        ph = torch.empty((9, int(n.value)), dtype=torch.float32, device=self.device)
        got = _U32()
        self.check(self.lib.pm_photon_file_read(self.h, os.fsencode(path), _ptr(ph), int(n.value), C.byref(got)))
        return ph[:, :got.value]

    def photon_file_write(self, path: str, photons9):
        ph = _f32_dev(photons9, self.device)
        assert ph.dim() == 2 and ph.shape[0] == 9
        self.check(self.lib.pm_photon_file_write(self.h, os.fsencode(path), _ptr(ph), ph.shape[1]))

    # -- scene / photon tracing ------------------------------------------------
    def upload_scene(self, scene, accel: bool = False) -> "DeviceScene":
        """Copy a scene to HBM; accel=True also builds its triangle BVHs."""
        ds = DeviceScene(self, scene)
        if accel:
            ds.build_accel()
        return ds

    def trace_photons(self, dscene: "DeviceScene", seed: int, n_global: int, n_caustic: int,
                      max_bounces: int = 10, max_paths: int = 1 << 40):
        import torch

        g = torch.empty((9, max(n_global, 1)), dtype=torch.float32, device=self.device)
        c = torch.empty((9, max(n_caustic, 1)), dtype=torch.float32, device=self.device)
        st = TraceStats()
        self.check(self.lib.pm_trace_photons(self.h, C.byref(dscene.rec), int(seed), int(n_global), int(n_caustic),
                                             int(max_bounces), int(max_paths), _ptr(g), _ptr(c), C.byref(st)))
        return g[:, :n_global], c[:, :n_caustic], (int(st.paths), int(st.n_global), int(st.n_caustic))

    def render_frame(self, dscene: "DeviceScene", grid_g: "Grid", grid_c, camera, max_depth=4, k_global=50,
                     r_global=0.05, k_caustic=50, r_caustic=0.05, flags=0, gather_flags=0, want_rgb8=True):
        """Render one frame; returns (rgb3 float32 [3, H*W], rgb8 uint8 [H*W*3] or None, hits)."""
        import torch

        cam = np.ascontiguousarray(camera)
        W, H = int(cam["width"].item()), int(cam["height"].item())
        rgb = torch.empty((3, W * H), dtype=torch.float32, device=self.device)
        rgb8 = torch.empty((W * H * 3,), dtype=torch.uint8, device=self.device) if want_rgb8 else None
        p = RenderParams(int(max_depth), int(k_global), float(r_global), int(k_caustic), float(r_caustic),
                         int(flags), int(gather_flags))
        st = FrameStats()
        self.check(self.lib.pm_render_frame(self.h, C.byref(dscene.rec), grid_g.h, grid_c.h if grid_c else None,
                                            cam.ctypes.data, C.byref(p), _ptr(rgb), _ptr(rgb8), C.byref(st)))
        return rgb, rgb8, int(st.hits)

    def pipeline(self, dscene: "DeviceScene", aabb, cell_global: float, cell_caustic: float, n_global: int,
                 n_caustic: int, camera, max_bounces: int = 10, max_paths: int = 1 << 40, plan_seed: int = 1,
                 max_depth: int = 4, k_global: int = 50, r_global: float = 0.05, k_caustic: int = 50,
                 r_caustic: float = 0.05, gather_flags: int = 0) -> "Pipeline":
        """pm_pipeline_create: a per-frame pipeline (trace, grid rebuilds,
        render) that pm_pipeline_frame enqueues without host synchronisation."""
        d = PipelineDesc()
        d.scene = C.addressof(dscene.rec)
        d.aabb = (C.c_float * 6)(*[float(v) for v in aabb])
        d.cell_global, d.cell_caustic = float(cell_global), float(cell_caustic)
        d.n_global, d.n_caustic = int(n_global), int(n_caustic)
        d.max_bounces, d.max_paths, d.plan_seed = int(max_bounces), int(max_paths), int(plan_seed)
        C.memmove(C.byref(d.camera), np.ascontiguousarray(camera).ctypes.data, C.sizeof(CameraRec))
        p = RenderParams(int(max_depth), int(k_global), float(r_global), int(k_caustic), float(r_caustic), 0,
                         int(gather_flags))
        C.memmove(C.byref(d.params), C.byref(p), C.sizeof(RenderParams))
        h = _P()
        self.check(self.lib.pm_pipeline_create(self.h, C.byref(d), C.byref(h)))
        cam = np.ascontiguousarray(camera)
        return Pipeline(self, h, dscene, int(cam["width"].item()) * int(cam["height"].item()))

    def frame_queries(self, dscene: "DeviceScene", camera, max_depth=4, k_global=50, r_global=0.05, k_caustic=50,
                      r_caustic=0.05):
        """A frame's gather queries (pm_frame_rays + pm_frame_copy_queries):
        (queries3, normals3) device tensors of the eye-ray hits."""
        import torch

        f = _P()
        self.check(self.lib.pm_frame_create(self.h, C.byref(f)))
        try:
            cam = np.ascontiguousarray(camera)
            p = RenderParams(int(max_depth), int(k_global), float(r_global), int(k_caustic), float(r_caustic), 0, 0)
            nq = _U32()
            self.check(self.lib.pm_frame_rays(self.h, f, C.byref(dscene.rec), cam.ctypes.data, C.byref(p),
                                              C.byref(nq)))
            q = torch.empty((3, max(nq.value, 1)), dtype=torch.float32, device=self.device)
            n = torch.empty_like(q)
            got = _U32()
            self.check(self.lib.pm_frame_copy_queries(self.h, f, _ptr(q), _ptr(n), nq.value, C.byref(got)))
            self.sync()
            return q[:, :nq.value], n[:, :nq.value]
        finally:
            self.lib.pm_frame_destroy(f)

    def render_frames_batched(self, frames, flags: int = 0, want_rgb8: bool = True):
        """Render many frames with their gathers in one pm_gather_batch:
        frames = [(dscene, grid_g, grid_c or None, camera, max_depth, k_global,
        r_global, k_caustic, r_caustic)]. Returns [(rgb3, rgb8, hits)]."""
        import torch

        handles, outs = [], []
        try:
            recs = (GatherJob * (2 * max(len(frames), 1)))()
            n_jobs = 0
            for (ds, gg, gc, camera, md, kg, rg, kc, rc) in frames:
                f = _P()
                self.check(self.lib.pm_frame_create(self.h, C.byref(f)))
                handles.append(f)
                cam = np.ascontiguousarray(camera)
                p = RenderParams(int(md), int(kg), float(rg), int(kc), float(rc), 0, 0)
                nq = _U32()
                self.check(self.lib.pm_frame_rays(self.h, f, C.byref(ds.rec), cam.ctypes.data, C.byref(p),
                                                  C.byref(nq)))
                jb = (GatherJob * 2)()
                nj = _U32()
                self.check(self.lib.pm_frame_gather_jobs(f, gg.h, gc.h if gc else None, C.cast(jb, C.c_void_p),
                                                         C.byref(nj)))
                for i in range(nj.value):
                    recs[n_jobs] = jb[i]
                    n_jobs += 1
            self.check(self.lib.pm_gather_batch(self.h, C.cast(recs, C.c_void_p), n_jobs, int(flags)))
            for f, (ds, gg, gc, camera, *_rest) in zip(handles, frames):
                cam = np.ascontiguousarray(camera)
                W, H = int(cam["width"].item()), int(cam["height"].item())
                rgb = torch.empty((3, W * H), dtype=torch.float32, device=self.device)
                rgb8 = torch.empty((W * H * 3,), dtype=torch.uint8, device=self.device) if want_rgb8 else None
                st = FrameStats()
                self.check(self.lib.pm_frame_finish(self.h, f, _ptr(rgb), _ptr(rgb8), C.byref(st)))
                outs.append((rgb, rgb8, int(st.hits)))
            torch.cuda.synchronize(self.device)
        finally:
            for f in handles:
                self.lib.pm_frame_destroy(f)
        return outs

    def server(self, budget_bytes: int = 0) -> "Server":
        return Server(self, budget_bytes)

    def box_stats(self, grid: "Grid", queries3, radius: float, reference: bool = False):
        """(sum C_q, sum P_q): the gather's visited cells / candidates, or with
        reference=True the reference's whole candidate box."""
        q = _f32_dev(queries3, self.device)
        c = C.c_uint64()
        p = C.c_uint64()
        self.check(self.lib.pm_grid_box_stats(self.h, grid.h, _ptr(q), q.shape[1], float(radius),
                                              BOX_REFERENCE if reference else 0, C.byref(c), C.byref(p)))
        return int(c.value), int(p.value)


@dataclass
class GridInfo:
    aabb: tuple
    cell_size: float
    dims: tuple
    layout: int
    n_cells: int
    n_photons: int


class Grid:
    def __init__(self, ctx: Context, h, keep=None):
        self.ctx = ctx
        self.h = h
        self._keep = keep
        d = GridDesc()
        ctx.check(ctx.lib.pm_grid_describe(h, C.byref(d)))
        self.desc = d
        self.info = GridInfo(tuple(d.aabb), d.cell_size, tuple(d.dims), d.layout, d.n_cells, d.n_photons)

    def rebuild(self, photons9, n: int | None = None):
        """pm_grid_rebuild: the same cells over a new photon block (at most
        the photon count the grid was built with)."""
        ph = _f32_dev(photons9, self.ctx.device)
        assert ph.dim() == 2 and ph.shape[0] == 9, "photons9 must be [9, n]"
        n = ph.shape[1] if n is None else int(n)
        self.ctx.check(self.ctx.lib.pm_grid_rebuild(self.ctx.h, self.h, _ptr(ph), n))
        self._keep = ph
        d = GridDesc()
        self.ctx.check(self.ctx.lib.pm_grid_describe(self.h, C.byref(d)))
        self.desc = d
        self.info = GridInfo(tuple(d.aabb), d.cell_size, tuple(d.dims), d.layout, d.n_cells, d.n_photons)
        return self

    def to_host(self):
        d = self.desc
        starts = np.empty(d.n_cells, np.uint32)
        sizes = np.empty(d.n_cells, np.uint32)
        photons = np.empty((9, d.n_photons), np.float32)
        perm = np.empty(d.n_photons, np.uint32)
        self.ctx.check(self.ctx.lib.pm_grid_copy_to_host(
            self.ctx.h, self.h, starts.ctypes.data, sizes.ctypes.data, photons.ctypes.data, perm.ctypes.data))
        return starts, sizes, photons, perm

    def close(self):
        if self.h:
            self.ctx.lib.pm_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pipeline:
    def __init__(self, ctx: Context, h, dscene, npix: int):
        self.ctx, self.h, self.dscene, self.npix = ctx, h, dscene, npix

    def frame(self, seed: int, rgb8=None, rgb3=None):
        """Enqueue one frame (pm_pipeline_frame) into the given device
        buffers (uint8 [H*W*3], float32 [3, H*W]; either may be None)."""
        self.ctx.check(self.ctx.lib.pm_pipeline_frame(self.h, int(seed), _ptr(rgb3), _ptr(rgb8)))

    def status(self) -> dict:
        info = PipelineInfo()
        self.ctx.check(self.ctx.lib.pm_pipeline_status(self.h, C.byref(info)))
        return {"scheduled_paths": int(info.scheduled_paths), "n_global": int(info.n_global),
                "n_caustic": int(info.n_caustic), "complete": bool(info.complete), "hits": int(info.hits)}

    def close(self):
        if self.h:
            self.ctx.lib.pm_pipeline_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    """A scene's triangles and normals resident in HBM plus its model table."""

    def __init__(self, ctx: Context, scene):
        import torch

        self.scene = scene
        self.tris = torch.from_numpy(np.ascontiguousarray(scene.tris9, np.float32)).to(ctx.device)
        self.nrms = torch.from_numpy(np.ascontiguousarray(scene.normals9, np.float32)).to(ctx.device)
        self.models = np.ascontiguousarray(scene.models)
        assert self.models.dtype.itemsize == C.sizeof(ModelRec)
        r = SceneRec()
        r.d_tris9 = self.tris.data_ptr()
        r.d_normals9 = self.nrms.data_ptr()
        r.n_tris = scene.tris9.shape[0]
        r.h_models = self.models.ctypes.data
        r.n_models = len(self.models)
        r.light_model = int(scene.light_model)
        for i in range(3):
            r.light_intensity[i] = float(scene.light_intensity[i])
        r.default_refr_index = float(scene.default_refr_index)
        r.accel = None
        self.rec = r
        self.ctx = ctx
        self.accel = None

    def build_accel(self):
        """Build (or rebuild) the per-model BVHs used by tracing and rendering."""
        self.drop_accel()
        a = _P()
        self.ctx.check(self.ctx.lib.pm_accel_build(self.ctx.h, C.byref(self.rec), C.byref(a)))
        self.accel = a
        self.rec.accel = a.value

    def drop_accel(self):
        if self.accel is not None:
            self.ctx.lib.pm_accel_destroy(self.accel)
            self.accel = None
            self.rec.accel = None

    def __del__(self):
        try:
            self.drop_accel()
        except Exception:
            pass


class KdTree:
    def __init__(self, ctx: Context, h, keep=None):
        self.ctx = ctx
        self.h = h
        self._keep = keep
        d = KdTreeDesc()
        ctx.check(ctx.lib.pm_kdtree_describe(h, C.byref(d)))
        self.desc = d

    def to_host(self):
        d = self.desc
        ni = np.empty(d.n_slots, np.uint32)
        npl = np.empty(d.n_slots, np.int8)
        self.ctx.check(self.ctx.lib.pm_kdtree_copy_to_host(self.ctx.h, self.h, ni.ctypes.data, npl.ctypes.data))
        return ni, npl

    def close(self):
        if self.h:
            self.ctx.lib.pm_kdtree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching render server (pm_server_*): submit jobs, step."""

    def __init__(self, ctx: Context, budget_bytes: int = 0):
        self.ctx = ctx
        self.h = _P()
        ctx.check(ctx.lib.pm_server_create(ctx.h, int(budget_bytes), C.byref(self.h)))
        self._keep = []

    def set_workers(self, n: int):
        """pm_server_set_workers: per-job stages on n worker contexts / streams."""
        self.ctx.check(self.ctx.lib.pm_server_set_workers(self.h, int(n)))

    def submit(self, dscene: "DeviceScene", aabb, cell_size, n_global, n_caustic, camera, frames, seed,
               max_depth=4, k_global=64, r_global=0.05, k_caustic=64, r_caustic=0.05, max_bounces=10,
               max_paths=1 << 40, gather_flags=0):
        """Queue a job; returns (job_id, rgb8 device tensor [frames, H*W*3])."""
        import torch

        cam = np.ascontiguousarray(camera)
        W, H = int(cam["width"].item()), int(cam["height"].item())
        out = torch.zeros((frames, W * H * 3), dtype=torch.uint8, device=self.ctx.device)
        d = JobDesc()
        d.scene = C.addressof(dscene.rec)
        for i in range(6):
            d.aabb[i] = float(aabb[i])
        d.cell_size = float(cell_size)
        d.n_global, d.n_caustic, d.max_bounces = int(n_global), int(n_caustic), int(max_bounces)
        d.max_paths, d.seed = int(max_paths), int(seed)
        C.memmove(C.byref(d.camera), cam.ctypes.data, C.sizeof(CameraRec))
        p = RenderParams(int(max_depth), int(k_global), float(r_global), int(k_caustic), float(r_caustic), 0,
                         int(gather_flags))
        C.memmove(C.byref(d.params), C.byref(p), C.sizeof(RenderParams))
        d.frames = int(frames)
        d.d_rgb8 = out.data_ptr()
        jid = C.c_uint64()
        self.ctx.check(self.ctx.lib.pm_server_submit(self.h, C.byref(d), C.byref(jid)))
        self._keep.append((dscene, out))
        return int(jid.value), out

    def step(self) -> int:
        n = _U32()
        self.ctx.check(self.ctx.lib.pm_server_step(self.h, C.byref(n)))
        return int(n.value)

    def info(self) -> dict:
        s = ServerInfo()
        self.ctx.check(self.ctx.lib.pm_server_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerInfo._fields_}

    def latencies(self) -> np.ndarray:
        n = C.c_uint64()
        self.ctx.check(self.ctx.lib.pm_server_latencies(self.h, None, 0, C.byref(n)))
        out = np.empty(n.value, np.float64)
        self.ctx.check(self.ctx.lib.pm_server_latencies(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def close(self):
        if self.h:
            self.ctx.lib.pm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
