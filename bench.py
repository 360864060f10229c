"""Benchmark of the photon gather hot path on one MI355X.

Workload: the gather microbenchmark -- 16M FP32 photons (70% uniform on the
unit-cube faces, 30% in 64 Gaussian clusters on the floor), 4M queries on the
faces, k = 64 nearest photons within r = 0.01. The hash-grid cell size
(r/4, r/3, 0.4r, r/2, 2r/3 or r) and the gather variant are chosen jointly by
the library's autotuner (pm_tune_gather); the neighbours are the same for
every choice. One step = rebuild the resident photon hash grid in place from
the photon block (keys -> radix sort -> cell sizes from the sorted runs ->
scan -> record gather) + the fused k-nearest gather and radiance estimate for
all 4M queries. Inputs are resident in HBM before the timed region.

Prints one JSON line: metric, value (queries/s), step time, the roofline of
the gather kernel and the CPU baseline, plus the per-frame, batched and
served forms.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-photon-mapping_amd"))

METRIC = "kNN photon queries/s (k=64) + photons traced/s, 1 GPU; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    """Host threads the CPU baselines use: OMP_NUM_THREADS, else all cores
    (at most 64)."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(threads, 64))


def native_oracle():
    """Builds the oracle with -march=native for this host (into a temporary
    directory) and points pyoracle at it, for the CPU baselines; falls back
    to the portable in-tree build. Returns the build used."""
    import subprocess
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    d = tempfile.mkdtemp(prefix="pm355_oracle_native_")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVE_DIR={d}"], check=True,
                       capture_output=True, timeout=300)
        pyoracle.LIB_PATH = os.path.join(d, "liboracle.so")
        pyoracle._lib = None
        return "g++ -O3 -march=native -ffp-contract=off -fopenmp, built on this host"
    except (OSError, subprocess.SubprocessError) as e:
        return f"g++ -O3 -ffp-contract=off -fopenmp, portable build (the -march=native build failed: {e})"


def insts_per_query(cs):
    """Wave instructions per query of the profiled gather launch (SQ_INSTS_*
    over SQ_WAVES; the launch runs one wave per query)."""
    m = (cs or {}).get("per_dispatch_median", {})
    w = m.get("SQ_WAVES")
    if not w:
        return None
    names = {"valu": "SQ_INSTS_VALU", "salu": "SQ_INSTS_SALU", "lds": "SQ_INSTS_LDS",
             "vmem_rd": "SQ_INSTS_VMEM_RD", "smem": "SQ_INSTS_SMEM"}
    return {k: m[c] / w for k, c in names.items() if c in m}


def valu_lane_util(cs):
    """Active lanes per VALU instruction over 64 (SQ_THREAD_CYCLES_VALU over
    64 x SQ_ACTIVE_INST_VALU)."""
    m = (cs or {}).get("per_dispatch_median", {})
    if m.get("SQ_ACTIVE_INST_VALU") and "SQ_THREAD_CYCLES_VALU" in m:
        return m["SQ_THREAD_CYCLES_VALU"] / (64.0 * m["SQ_ACTIVE_INST_VALU"])
    return None


def cpu_baseline(ph, q, qn, aabb, r, k, build, run_s=3.0, runs=5):
    """Oracle (C++ restatement of the reference CPU path) on the host cores:
    grid build once, then kNN + radiance over a seeded sample of the queries
    sized for about run_s seconds, timed `runs` times; the median counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import statistics

    import numpy as np

    import pyoracle

    threads = host_threads()
    t0 = time.perf_counter()
    og = pyoracle.OracleGrid(ph, aabb, cell_size=r)
    t_build = time.perf_counter() - t0

    def run(sel):
        t0 = time.perf_counter()
        oi, od, oc = og.knn(q[:, sel], r, k, threads=threads)
        pyoracle.radiance(oi, od, oc, og.photons, qn[:, sel], k)
        return time.perf_counter() - t0

    pilot = np.random.default_rng(0).choice(q.shape[1], 20_000, replace=False)
    dt = run(pilot)
    n = int(min(q.shape[1], max(20_000, 20_000 * run_s / max(dt, 1e-3))))
    sel = np.random.default_rng(1).choice(q.shape[1], n, replace=False)
    ts = [run(sel) for _ in range(runs)]
    return {
        "value": n / statistics.median(ts),
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} of the 4M queries (seeded sample), kNN k={k} + radiance on the oracle grid at cell = r, "
                  f"median of {runs} runs ({', '.join(f'{t:.2f}' for t in ts)} s); oracle {build}; oracle grid "
                  f"build of the 16M photons took {t_build:.2f} s on 1 thread (not in value)",
    }


def cpu_stages(threads, build):
    """Per-stage timings of the oracle on the host cores (the reference CPU
    path restated), each the median of five runs on identical inputs:
    photon tracing of the Cornell box (1 thread and all threads), a
    2M-photon grid build and kd-tree build (1 thread), the kNN gather at
    k = 64 (1 thread and all threads), the radiance estimate over its result
    (1 thread), and a 320x240 frame (100k photons, k = 50, r = 0.05) on all
    threads. The units say what was counted."""
    import statistics

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    from pm355.scenes import cornell_box, cornell_camera
    from pm355.workloads import gather_microbench

    def med(fn, reps=5):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    out = {"threads_all": threads, "nproc": os.cpu_count(), "runs_per_stage": 5, "oracle_build": build}
    try:
        info = {}
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ":" in ln:
                    key, val = ln.split(":", 1)
                    info.setdefault(key.strip(), val.strip())
        flags = set(info.get("flags", "").split())
        out["cpu_model"] = info.get("model name", "")
        out["cpu_family_model_stepping"] = [info.get("cpu family"), info.get("model"), info.get("stepping")]
        out["cpu_simd"] = sorted(f for f in flags if f in ("avx2", "fma", "avx512f", "avx512vl", "avx512_bf16"))
    except OSError:
        pass
    sc = cornell_box()
    res = {}

    def trace():
        res["trace"] = pyoracle.trace_photons(sc, 7, 100_000, 10_000, 8)

    pyoracle.set_threads(1)
    dt = med(trace)
    pyoracle.set_threads(threads)
    st = res["trace"][2]
    out["trace_photons_per_s_1thread"] = (st[1] + st[2]) / dt
    out["trace_photons_per_s_all"] = (st[1] + st[2]) / med(trace)
    pyoracle.set_threads(0)
    ph, q, qn, aabb = gather_microbench(seed=3, n_photons=2_000_000, n_queries=20_000)
    out["grid_build_2M_ms_1thread"] = med(lambda: pyoracle.OracleGrid(ph, aabb, cell_size=0.01)) * 1e3
    og = pyoracle.OracleGrid(ph, aabb, cell_size=0.01)
    out["kdtree_build_2M_ms_1thread"] = med(lambda: pyoracle.kdtree_build(ph[0:3])) * 1e3
    for name, th in (("1thread", 1), ("all", threads)):
        nq = 5_000 if th == 1 else 20_000
        out[f"knn_queries_per_s_{name}"] = nq / med(lambda: og.knn(q[:, :nq], 0.01, 64, threads=th))
    oi, od, oc = og.knn(q[:, :5_000], 0.01, 64, threads=threads)
    out["radiance_queries_per_s_1thread"] = 5_000 / med(
        lambda: pyoracle.radiance(oi, od, oc, og.photons, qn[:, :5_000], 64))
    g, c, _ = res["trace"]
    gg = pyoracle.OracleGrid(g, sc.aabb(), cell_size=0.05)
    gc = pyoracle.OracleGrid(c, sc.aabb(), cell_size=0.05)
    cam = cornell_camera(320, 240)
    out["frame_320x240_ms_all"] = med(
        lambda: pyoracle.render(sc, gg, gc, cam, 4, 50, 0.05, 50, 0.05, threads=threads)) * 1e3
    return out


def frame_bench(ctx, n_frames):
    """The per-frame form on the benchmark's Cornell box: 2M global + 500k
    caustic photons traced, both hash grids (cell = r) and both kd-trees
    built, then a 1280x720 frame with k = 100 within r = 0.02, max depth 4.
    The render reads the grids only, so the kd-tree builds run on a second
    context and stream (default priority) while the grids and the render run
    on the first; a frame ends when both streams are done. Reports photons
    traced/s and p50/p99 frame latency (wall clock around a synchronized
    frame) with the per-stage medians (HIP events), and the same frame with
    every stage on one stream."""
    import numpy as np
    import torch

    import pm355
    from pm355.scenes import cornell_box, cornell_camera

    scene = cornell_box()
    ds = ctx.upload_scene(scene, accel=True)
    aabb = scene.aabb()
    cam = cornell_camera(1280, 720)
    stream = torch.cuda.current_stream(ctx.device)
    kd_stream = torch.cuda.Stream(ctx.device)
    with torch.cuda.stream(kd_stream):
        ctx_kd = pm355.Context(ctx.device.index)

    def frame(seed, overlap, cell_g=0.02, cell_c=0.02, gflags=0):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        kev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        g, c, st = ctx.trace_photons(ds, seed, 2_000_000, 500_000, 10, max_paths=200_000_000)
        ev[1].record(stream)
        if overlap:
            kd_stream.wait_event(ev[1])
            with torch.cuda.stream(kd_stream):
                kev[0].record(kd_stream)
                kg, kc = ctx_kd.kdtree_build(g), ctx_kd.kdtree_build(c)
                kev[1].record(kd_stream)
        gg = ctx.grid_build(g, aabb, cell_size=cell_g)
        gc = ctx.grid_build(c, aabb, cell_size=cell_c)
        ev[2].record(stream)
        if not overlap:
            kev[0].record(stream)
            kg, kc = ctx.kdtree_build(g), ctx.kdtree_build(c)
            kev[1].record(stream)
        ev[3].record(stream)
        rgb, rgb8, hits = ctx.render_frame(ds, gg, gc, cam, 4, 100, 0.02, 100, 0.02, gather_flags=gflags)
        ev[4].record(stream)
        if overlap:
            stream.wait_stream(kd_stream)
        torch.cuda.synchronize()
        for h in (gg, gc, kg, kc):
            h.close()
        return st, [ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), kev[0].elapsed_time(kev[1]),
                    ev[3].elapsed_time(ev[4])]

    out = {}
    for overlap in (True, False):
        frame(1, overlap)
        lat, stages, stored = [], [], 0
        for i in range(n_frames):
            t0 = time.perf_counter()
            st, sm = frame(100 + i, overlap)
            lat.append((time.perf_counter() - t0) * 1e3)
            stages.append(sm)
            stored = st[1] + st[2]
        stages = np.median(np.array(stages), 0)
        out[overlap] = (lat, stages, stored)
    # Tuned: the library's autotuner on one frame's maps and the frame's own
    # gather queries (its eye-ray hits), the global and the caustic map each;
    # both grids at their tuned cell sizes and the render's gathers flagged
    # PM_KNN_TUNED.
    g, c, _ = ctx.trace_photons(ds, 7, 2_000_000, 500_000, 10, max_paths=200_000_000)
    tq, tn = ctx.frame_queries(ds, cam, 4, 100, 0.02, 100, 0.02)
    tuning = ctx.tune_gather(g, aabb, tq, tn, 0.02, 100)
    tuning_c = ctx.tune_gather(c, aabb, tq, tn, 0.02, 100)
    del g, c, tq, tn
    cg = 0.02 * ctx.tuning_lookup(2_000_000, aabb, 0.02, 100)["cell_over_r"]
    cc = 0.02 * ctx.tuning_lookup(500_000, aabb, 0.02, 100)["cell_over_r"]
    frame(1, True, cg, cc, pm355.KNN_TUNED)
    tlat, tstages = [], []
    for i in range(n_frames):
        t0 = time.perf_counter()
        _, sm = frame(100 + i, True, cg, cc, pm355.KNN_TUNED)
        tlat.append((time.perf_counter() - t0) * 1e3)
        tstages.append(sm)
    tstages = np.median(np.array(tstages), 0)
    ctx_kd.close()
    # Pipelined: the same frame without the kd-trees (the render reads the
    # grids only) as one pm_pipeline frame -- trace, both grids rebuilt in
    # place, eye rays, both gathers, fold -- enqueued with every count on the
    # device, and the frame captured into a HIP graph and replayed. Its
    # context (one pipeline per context) runs on a stream of its own, as
    # graph capture needs; the frame is checked against the eager calls.
    pstream = torch.cuda.Stream(ctx.device)
    npix = 1280 * 720
    with torch.cuda.stream(pstream):
        pctx = pm355.Context(ctx.device.index)
        pl = pctx.pipeline(ds, aabb, 0.02, 0.02, 2_000_000, 500_000, cam, max_bounces=10, max_paths=200_000_000,
                           plan_seed=1, max_depth=4, k_global=100, r_global=0.02, k_caustic=100, r_caustic=0.02)
        p8 = torch.zeros(npix * 3, dtype=torch.uint8, device=ctx.device)
        p3 = torch.zeros((3, npix), dtype=torch.float32, device=ctx.device)
        plat = []
        for i in range(n_frames + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pl.frame(100 + i, p8, p3)
            pstream.synchronize()
            if i > 0:
                plat.append((time.perf_counter() - t0) * 1e3)
        pinfo = pl.status()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=pstream):
            rc = pctx.lib.pm_pipeline_frame(pl.h, 7, pm355._ptr(p3), pm355._ptr(p8))
        pctx.check(rc)
        glat = []
        for i in range(n_frames + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            graph.replay()
            pstream.synchronize()
            if i > 0:
                glat.append((time.perf_counter() - t0) * 1e3)
        g8 = p8.clone()
    g, c, _ = ctx.trace_photons(ds, 7, 2_000_000, 500_000, 10, max_paths=200_000_000)
    gg, gc = ctx.grid_build(g, aabb, cell_size=0.02), ctx.grid_build(c, aabb, cell_size=0.02)
    _, e8, _ = ctx.render_frame(ds, gg, gc, cam, 4, 100, 0.02, 100, 0.02)
    torch.cuda.synchronize()
    pipelined = {"p50_ms": float(np.percentile(plat, 50)), "p99_ms": float(np.percentile(plat, 99)),
                 "graph_replay_p50_ms": float(np.percentile(glat, 50)),
                 "graph_replay_p99_ms": float(np.percentile(glat, 99)),
                 "maps_complete": bool(pinfo["complete"]), "frame_equals_eager": bool(torch.equal(g8, e8)),
                 "note": "no kd-trees (the render reads the grids only); compare one_stream_stage_ms trace + "
                         "grid_builds + render"}
    for h in (gg, gc):
        h.close()
    del graph, g, c
    pl.close()
    pctx.close()
    lat, stages, stored = out[True]
    slat = out[False][0]
    valu = None
    vpath = os.path.join(ROOT, "profiles", "cornell_c2_trace_valu.json")
    if os.path.exists(vpath):
        # FP32 VALU operations of one C2 photon trace, counted by rocprofv3
        # (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F32 x 64 lanes, FMA x 2), and
        # whether the HIP sources changed since they were counted
        from pm355 import profiling

        v = json.load(open(vpath))
        tflops = v["fp32_flop_per_trace"] / (stages[0] * 1e-3) / 1e12
        valu = {"fp32_tflops": tflops, "peak_tflops": 157.3, "frac": tflops / 157.3,
                "source": "profiles/cornell_c2_trace_valu.json",
                "stale": v.get("source_hash") != profiling.source_hash(v.get("source_files"))}
    return {
        "workload": "Cornell box (10 wall + 2x960 sphere + 2 light triangles, per-model BVHs), 2M global + 500k "
                    "caustic photons (<=10 bounces), grids cell=r and kd-trees built per frame, 1280x720, "
                    "k=100, r=0.02, max depth 4; kd-tree builds on a second stream beside grids + render",
        "frames": n_frames,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "one_stream_p50_ms": float(np.percentile(slat, 50)),
        "one_stream_p99_ms": float(np.percentile(slat, 99)),
        "photons_traced_per_s": stored / (stages[0] * 1e-3),
        # tracing writes 36 B per stored photon (position, power, direction)
        "trace_hbm_frac": 36 * stored / (stages[0] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "stage_ms": {"trace": stages[0], "grid_builds": stages[1], "kdtree_builds": stages[2],
                     "render": stages[3]},
        "one_stream_stage_ms": {"trace": out[False][1][0], "grid_builds": out[False][1][1],
                                "kdtree_builds": out[False][1][2], "render": out[False][1][3]},
        "trace_valu": valu,
        "pipelined": pipelined,
        "tuned": {"p50_ms": float(np.percentile(tlat, 50)), "p99_ms": float(np.percentile(tlat, 99)),
                  "stage_ms": {"trace": tstages[0], "grid_builds": tstages[1], "kdtree_builds": tstages[2],
                               "render": tstages[3]},
                  "cell_over_r": {"global": cg / 0.02, "caustic": cc / 0.02},
                  "variant": {m: {v: n_ for n_, v in pm355.KNN_VARIANTS.items()}.get(t_["flags"])
                              for m, t_ in (("global", tuning), ("caustic", tuning_c))},
                  "p50_speedup_vs_default": float(np.percentile(lat, 50)) / float(np.percentile(tlat, 50))},
    }


def reference_frame_bench(ctx, n_frames, cpu_threads=0):
    """The reference application's own frame: its CornellBox-Sphere scene
    (2188 triangles, per-model BVHs), 1M global + 100k caustic photons
    (<= 10 bounces), both grids with the reference's cell sizing, then a
    1024x768 frame with k = 2000 within r = 0.1 (global) and k = 400 within
    r = 0.2 (caustic), max depth 4. Reports p50/p99 frame latency and stage
    medians (HIP events). With cpu_threads > 0 the oracle is timed on a
    bounded sample of the same frame: a 1/10 photon trace (100k + 10k,
    scaled x10) and a 256x192 render over the full-size maps (1/16 of the
    pixels, scaled x16)."""
    import numpy as np
    import torch

    from pm355.objload import load_obj
    from pm355.scenes import REFERENCE_FRAME as RF
    from pm355.scenes import reference_camera

    scene = load_obj(os.path.join(ROOT, "tests", "golden", "cornell_box_sphere"), "CornellBox-Sphere.obj")
    ds = ctx.upload_scene(scene, accel=True)
    aabb = scene.aabb()
    cam = reference_camera(RF["width"], RF["height"])
    stream = torch.cuda.current_stream(ctx.device)
    args = (RF["max_depth"], RF["k_global"], RF["r_global"], RF["k_caustic"], RF["r_caustic"])

    def frame(seed, keep=False, cell_over_r=0.0):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        g, c, st = ctx.trace_photons(ds, seed, RF["n_global"], RF["n_caustic"], RF["max_bounces"])
        ev[1].record(stream)
        gg = ctx.grid_build(g, aabb, cell_size=cell_over_r * RF["r_global"])
        gc = ctx.grid_build(c, aabb, cell_size=cell_over_r * RF["r_caustic"])
        ev[2].record(stream)
        rgb, rgb8, hits = ctx.render_frame(ds, gg, gc, cam, *args)
        ev[3].record(stream)
        torch.cuda.synchronize()
        gg.close()
        gc.close()
        sm = [ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), ev[2].elapsed_time(ev[3])]
        return st, hits, sm, ((g, c) if keep else None)

    _, _, _, maps = frame(1, keep=True)
    lat, stages = [], []
    for i in range(n_frames):
        t0 = time.perf_counter()
        st, hits, sm, _ = frame(200 + i)
        lat.append((time.perf_counter() - t0) * 1e3)
        stages.append(sm)
    stages = np.median(np.array(stages), 0)
    # the same frames with grid cells of r/2 (the gather's tuned cell size;
    # neighbour sets and images are unchanged by the cell size)
    frame(1, cell_over_r=0.5)
    tlat, tstages = [], []
    for i in range(n_frames):
        t0 = time.perf_counter()
        _, _, sm, _ = frame(200 + i, cell_over_r=0.5)
        tlat.append((time.perf_counter() - t0) * 1e3)
        tstages.append(sm)
    tstages = np.median(np.array(tstages), 0)
    out = {
        "workload": "reference frame: CornellBox-Sphere OBJ/MTL (2188 triangles, BVHs), 1M global + 100k caustic "
                    "photons (<=10 bounces), grids with the reference sizing, 1024x768, k=2000 r=0.1 global, "
                    "k=400 r=0.2 caustic, max depth 4",
        "frames": n_frames,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "stage_ms": {"trace": stages[0], "grid_builds": stages[1], "render": stages[2]},
        "eye_ray_hits": int(hits),
        "gather_queries_per_s": 2 * int(hits) / (stages[2] * 1e-3),
        "photons_traced_per_s": (int(st[1]) + int(st[2])) / (stages[0] * 1e-3),
        "cell_half_r": {"p50_ms": float(np.percentile(tlat, 50)), "p99_ms": float(np.percentile(tlat, 99)),
                        "stage_ms": {"trace": tstages[0], "grid_builds": tstages[1], "render": tstages[2]}},
    }
    if cpu_threads > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        O.set_threads(cpu_threads)
        t0 = time.perf_counter()
        O.trace_photons(scene, 1, RF["n_global"] // 10, RF["n_caustic"] // 10, RF["max_bounces"])
        trace_s = (time.perf_counter() - t0) * 10
        O.set_threads(0)
        gh, ch = (m.cpu().numpy() for m in maps)
        t0 = time.perf_counter()
        og, oc = O.OracleGrid(gh, aabb), O.OracleGrid(ch, aabb)
        grid_s = time.perf_counter() - t0
        small = reference_camera(RF["width"] // 4, RF["height"] // 4)
        t0 = time.perf_counter()
        O.render(scene, og, oc, small, *args, threads=cpu_threads)
        render_s = (time.perf_counter() - t0) * 16
        out["cpu_oracle"] = {"frame_ms_estimate": (trace_s + grid_s + render_s) * 1e3,
                             "stage_ms_estimate": {"trace": trace_s * 1e3, "grid_builds": grid_s * 1e3,
                                                   "render": render_s * 1e3},
                             "threads": {"trace": cpu_threads, "grid_builds": 1, "render": cpu_threads},
                             "sample": "trace 100k + 10k photons (x10), render 256x192 over the full maps (x16)"}
    return out


def batch_bench(ctx, n_jobs, reps=2):
    """Plain per-job form vs the batched persistent form on the many-job
    gather mix (photon counts log-uniform 10k-4M, one query per pixel of a
    128x128 / 256x256 / 640x480 / 1280x720 frame, k = 64, r in [0.01, 0.05],
    cell = r). Both forms gather radiance for every job's queries in cell
    order with the same kernel code and give identical results; the plain
    form makes one library call per job, the batched form one call (one
    processing-order sort and one persistent launch) for all jobs."""
    import torch

    import pm355
    from pm355.workloads import batch_mix, gather_inputs_device

    mix = batch_mix(n_jobs=n_jobs)
    jobs, grids, nq_total, n_ph = [], [], 0, 0
    for j in mix:
        ph, q, qn, aabb = gather_inputs_device(j["photons"], j["queries"], j["seed"], ctx.device)
        g = ctx.grid_build(ph, aabb, cell_size=j["radius"])
        del ph
        grids.append(g)
        jobs.append((g, q, qn, j["radius"], j["k"]))
        nq_total += j["queries"]
        n_ph += j["photons"]
    torch.cuda.synchronize()
    plain_flags = pm355.KNN_SORT_QUERIES

    def plain():
        for g, q, qn, r, k in jobs:
            ctx.gather_radiance(g, q, qn, r, k, flags=plain_flags)

    def batched():
        ctx.gather_batch(jobs, flags=pm355.KNN_SORT_QUERIES, want_count=False)

    def batched_queue():
        ctx.gather_batch(jobs, flags=pm355.KNN_SORT_QUERIES | pm355.KNN_BATCH_QUEUE, want_count=False)

    # the small jobs alone (128x128 and 256x256 frames): the launch-bound
    # regime, where one launch for many jobs pays most
    small = [j for j in jobs if j[1].shape[1] <= 256 * 256]
    nq_small = sum(j[1].shape[1] for j in small)

    def plain_small():
        for g, q, qn, r, k in small:
            ctx.gather_radiance(g, q, qn, r, k, flags=plain_flags)

    def batched_small():
        ctx.gather_batch(small, flags=pm355.KNN_SORT_QUERIES, want_count=False)

    res = {}
    for name, fn in (("plain", plain), ("batched", batched), ("batched_queue", batched_queue),
                     ("plain_small", plain_small), ("batched_small", batched_small)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / reps
    for g in grids:
        g.close()
    grids.clear()

    # Tuned: the library's autotuner on four jobs spread over the mix's photon
    # densities (n r^3), every job's grid at its density's tuned cell size and
    # its gathers flagged PM_KNN_TUNED (the nearest tuned density's variant).
    dens = sorted(range(len(mix)), key=lambda i: mix[i]["photons"] * mix[i]["radius"] ** 3)
    picks = sorted({dens[int(f * (len(dens) - 1))] for f in (0.2, 0.5, 0.8, 1.0)})
    tunings = []
    for i in picks:
        j = mix[i]
        ph, q, qn, aabb = gather_inputs_device(j["photons"], j["queries"], j["seed"], ctx.device)
        t = ctx.tune_gather(ph, aabb, q, qn, j["radius"], j["k"])
        tunings.append({"photons": j["photons"], "queries": j["queries"], "radius": j["radius"],
                        "cell_over_r": t["cell_over_r"], "flags": t["flags"], "ms": t["ms"],
                        "ms_default": t["ms_default"]})
        del ph, q, qn
    tjobs = []
    for j, (_, q, qn, r, k) in zip(mix, jobs):
        ph, _, _, aabb = gather_inputs_device(j["photons"], 1, j["seed"], ctx.device)
        cell = r * ctx.tuning_lookup(j["photons"], aabb, r, k)["cell_over_r"]
        g = ctx.grid_build(ph, aabb, cell_size=cell)
        del ph
        grids.append(g)
        tjobs.append((g, q, qn, r, k, pm355.KNN_TUNED))
    torch.cuda.synchronize()

    def plain_tuned():
        for g, q, qn, r, k, f in tjobs:
            ctx.gather_radiance(g, q, qn, r, k, flags=plain_flags | f)

    def batched_tuned():
        ctx.gather_batch(tjobs, flags=pm355.KNN_SORT_QUERIES, want_count=False)

    for name, fn in (("plain_tuned", plain_tuned), ("batched_tuned", batched_tuned)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / reps
    for g in grids:
        g.close()
    return {
        "workload": f"{n_jobs}-job gather mix: photons log-uniform 10k-4M per job ({n_ph} total), one query per "
                    f"pixel of 128^2/256^2/640x480/1280x720 frames ({nq_total} total), k=64, r U[0.01,0.05], "
                    f"cell=r; fused gather+radiance, cell-ordered",
        "plain_ms": res["plain"] * 1e3,
        "batched_ms": res["batched"] * 1e3,
        "batched_queue_ms": res["batched_queue"] * 1e3,
        "plain_queries_per_s": nq_total / res["plain"],
        "batched_queries_per_s": nq_total / res["batched"],
        "speedup": res["plain"] / res["batched"],
        "small_jobs": {"jobs": len(small), "queries": nq_small, "plain_ms": res["plain_small"] * 1e3,
                       "batched_ms": res["batched_small"] * 1e3,
                       "speedup": res["plain_small"] / res["batched_small"]},
        "tuned": {"plain_ms": res["plain_tuned"] * 1e3, "batched_ms": res["batched_tuned"] * 1e3,
                  "plain_speedup_vs_default": res["plain"] / res["plain_tuned"],
                  "batched_speedup_vs_default": res["batched"] / res["batched_tuned"],
                  "tunings": tunings,
                  "note": "grids at each job's tuned cell size, gathers flagged PM_KNN_TUNED; the batched launch "
                          "takes each job's inner-box choice, its kernel is the 512-key stream for all jobs"},
    }


def frames_batch_bench(ctx, n_scenes=256, reps=2, seed=7):
    """The batched form on whole frames of the many-scene mix: n_scenes
    scenes of the batched distribution (unit box + 1-8 spheres of 200-20k
    triangles, BVHs; diffuse / mirror / glass), global photons log-uniform in
    10k-4M (+5% caustic when the scene has a specular or refractive model;
    at most 3 paths per requested photon), canvases of 128^2 / 256^2 /
    640x480 / 1280x720, k = 64, r U[0.01, 0.05], cells of r, depth 4.

    plain: one scene's frame at a time, as the reference renders -- trace,
    both grid builds, pm_render_frame (its own gather launch).
    batched: every scene's trace and grids, then every frame's eye rays, the
    gathers of all frames in one pm_gather_batch, and each frame's fold.
    served: the scenes as one-frame jobs of the continuous-batching server
    (one step renders them all), on its context alone or with three worker
    contexts whose streams run the per-job trace / grid / eye-ray stages
    concurrently. Every form renders the same frames (checked)."""
    import numpy as np
    import torch

    from pm355.scenes import camera, random_scene

    # the earlier sections' cached device blocks go back first: the server's
    # default budget is a share of the memory free when it is created
    ctx.trim()
    torch.cuda.empty_cache()
    rng = np.random.default_rng(seed)
    res_set = [(128, 128), (256, 256), (640, 480), (1280, 720)]
    jobs = []
    for i in range(n_scenes):
        sc = random_scene(1000 + i)
        spec = any(float(m["opaque"]) != 1.0 or bool(np.any(m["ks"] != 0.0))
                   for m, nm in zip(sc.models, sc.names) if nm.startswith("sphere"))
        ng = int(np.exp(rng.uniform(np.log(1e4), np.log(4e6))))
        w, h = res_set[int(rng.integers(0, len(res_set)))]
        r = float(rng.uniform(0.01, 0.05))
        jobs.append(dict(scene=sc, aabb=sc.aabb(), ng=ng, nc=ng // 20 if spec else 0, r=r,
                         seed=int(rng.integers(1, 1 << 62)),
                         cam=camera((0.5, 0.5, 1.85), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 60.0, w, h)))
    log(f"frames mix: {n_scenes} scenes drawn")
    for i, j in enumerate(jobs):
        j["ds"] = ctx.upload_scene(j["scene"], accel=True)
    torch.cuda.synchronize()
    log("frames mix: scenes resident")
    n_ph = [0]
    n_q = [0]

    def maps(j):
        g, c, st = ctx.trace_photons(j["ds"], j["seed"], j["ng"], j["nc"], 10, max_paths=3 * (j["ng"] + j["nc"]))
        gg = ctx.grid_build(g[:, :st[1]], j["aabb"], cell_size=j["r"])
        gc = ctx.grid_build(c[:, :st[2]], j["aabb"], cell_size=j["r"]) if st[2] > 0 else None
        n_ph[0] += st[1] + st[2]
        return gg, gc

    def plain():
        outs = []
        for j in jobs:
            gg, gc = maps(j)
            _, rgb8, hits = ctx.render_frame(j["ds"], gg, gc, j["cam"], 4, 64, j["r"], 64, j["r"])
            n_q[0] += hits * (2 if gc else 1)
            outs.append(rgb8)
            gg.close()
            if gc:
                gc.close()
        return outs

    def batched():
        grids = [maps(j) for j in jobs]
        res = ctx.render_frames_batched([(j["ds"], gg, gc, j["cam"], 4, 64, j["r"], 64, j["r"])
                                         for j, (gg, gc) in zip(jobs, grids)])
        for gg, gc in grids:
            gg.close()
            if gc:
                gc.close()
        return [r8 for _, r8, _ in res]

    # The server's frame f of a job is seeded seed + (f + 1) * step: a
    # one-frame job submitted with seed - step renders the plain frame.
    seed_step = 0x9E3779B97F4A7C15

    servers = []
    # every server gets the same admission budget, a share of the device's
    # memory (the default, a share of the memory free at creation, depends on
    # what earlier sections left allocated; admission counts each job's
    # worst-case footprint, far above what its frame allocates)
    budget = int(0.8 * torch.cuda.get_device_properties(ctx.device).total_memory)
    steps_taken = {}

    def served(workers):
        # one server (and its worker contexts, with their block caches) for
        # every repetition of the form, as a serving process keeps them
        def fn():
            if not servers or servers[-1][0] != workers:
                # the previous form's server (and the blocks it cached) goes
                # first: each form's default budget sees the same free memory
                while servers:
                    servers.pop()[1].close()
                ctx.trim()
                torch.cuda.empty_cache()
                srv = ctx.server(budget)
                if workers:
                    srv.set_workers(workers)
                servers.append((workers, srv))
            srv = servers[-1][1]
            done0 = srv.info()["finished"]
            outs = []
            for j in jobs:
                _, o = srv.submit(j["ds"], j["aabb"], j["r"], j["ng"], j["nc"], j["cam"], 1,
                                  (j["seed"] - seed_step) & ((1 << 64) - 1), max_depth=4, k_global=64,
                                  r_global=j["r"], k_caustic=64, r_caustic=j["r"],
                                  max_paths=3 * (j["ng"] + j["nc"]))
                outs.append(o)
            steps = 0
            while srv.info()["finished"] < done0 + len(jobs):
                srv.step()
                steps += 1
            steps_taken[workers] = steps
            torch.cuda.synchronize()
            return [o[0] for o in outs]
        return fn

    out = {}
    first = {}
    for name, fn in (("plain", plain), ("batched", batched), ("served", served(0)),
                     ("served_workers3", served(3))):
        log(f"frames mix: {name}")
        first[name] = fn()
        torch.cuda.synchronize()
        n_ph[0] = n_q[0] = 0
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / reps
        out[name + "_photons"] = n_ph[0] // reps
        if name == "plain":
            out["queries"] = n_q[0] // reps
    same = {name: all(torch.equal(a, b) for a, b in zip(first["plain"], first[name]))
            for name in ("batched", "served", "served_workers3")}
    for _, srv in servers:
        srv.close()
    jobs.clear()
    return {
        "workload": f"{n_scenes} scenes of the batched distribution (unit box + 1-8 spheres of 200-20k triangles, "
                    "BVHs; diffuse/mirror/glass), global photons log-uniform 10k-4M (+5% caustic, <= 3 paths per "
                    "photon), 128^2/256^2/640x480/1280x720 canvases, k=64, r U[0.01,0.05], cell=r, depth 4; one "
                    "frame per scene (trace, grids, eye rays, gathers, fold)",
        "plain_ms": out["plain"] * 1e3,
        "batched_ms": out["batched"] * 1e3,
        "served_ms": out["served"] * 1e3,
        "served_workers3_ms": out["served_workers3"] * 1e3,
        "speedup": out["plain"] / out["batched"],
        "speedup_served_workers3": out["plain"] / out["served_workers3"],
        "photons_traced": out["plain_photons"],
        "gather_queries": out["queries"],
        "frames_identical": same,
        "server_budget_bytes": budget,
        "served_steps": {"served": steps_taken.get(0), "served_workers3": steps_taken.get(3)},
        "note": "batched = every scene's trace and grids, then one pm_gather_batch for every frame's gathers; "
                "served = all 256 scenes as one-frame jobs of pm_server (one step: per-job trace, grids and eye "
                "rays, one shared gather batch, per-frame fold), on the server's context or with three worker "
                "contexts running the per-job stages on their own streams",
    }


def serve_bench(ctx, rates=(10, 50, 200), n_jobs=50, seed=11,
                configs=(("default", 0, 0), ("tuned", 0, 1), ("tuned_workers3", 3, 1))):
    """The served form: Poisson arrivals of render jobs (scenes of the batched
    distribution, resident with their BVHs; 50k-8M photons, 5% caustic;
    frames of 128x128 up to 1920x1080; 1-30 frames each; k = 64,
    r in [0.01, 0.05], cell = r) into the continuous-batching server.
    Reports per arrival rate and server configuration the p50 / p99 frame
    latency (eligible -> completed) and gather queries/s over the run, with
    the jobs as given (cell = r) and flagged PM_KNN_TUNED (grid cell and
    gather variant from the context's tunings, the nearest tuned photon
    density), for the same arrivals."""
    import numpy as np
    import torch

    from pm355.scenes import camera, random_scene

    ctx.trim()
    torch.cuda.empty_cache()
    scenes = [random_scene(100 + i) for i in range(6)]
    boxes = [sc.aabb() for sc in scenes]
    dss = [ctx.upload_scene(sc, accel=True) for sc in scenes]
    res_set = [(128, 128), (256, 256), (640, 480), (1280, 720), (1920, 1080)]
    import pm355

    out = {}
    for rate, (cname, workers, tuned) in [(r_, c_) for r_ in rates for c_ in configs]:
        rng = np.random.default_rng(seed + rate)
        jobs = []
        t = 0.0
        for j in range(n_jobs):
            t += rng.exponential(1.0 / rate)
            s = int(rng.integers(0, len(scenes)))
            w, h = res_set[int(rng.integers(0, len(res_set)))]
            ng = int(np.exp(rng.uniform(np.log(5e4), np.log(8e6))))
            r = float(rng.uniform(0.01, 0.05))
            jobs.append(dict(at=t, scene=s, w=w, h=h, ng=ng, nc=ng // 20, frames=int(rng.integers(1, 31)), r=r,
                             seed=int(rng.integers(1, 1 << 62))))
        # every configuration's server gets the same admission budget, a share
        # of the device's memory (not of what earlier sections left free)
        ctx.trim()
        torch.cuda.empty_cache()
        srv = ctx.server(int(0.8 * torch.cuda.get_device_properties(ctx.device).total_memory))
        if workers:
            srv.set_workers(workers)
        log(f"served: {rate} jobs/s, {cname}")
        t0 = time.perf_counter()
        nxt = 0
        while nxt < len(jobs) or srv.info()["finished"] < len(jobs):
            now = time.perf_counter() - t0
            while nxt < len(jobs) and jobs[nxt]["at"] <= now:
                jb = jobs[nxt]
                sc = scenes[jb["scene"]]
                cam = camera((0.5, 0.5, 1.85), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 60.0, jb["w"], jb["h"])
                srv.submit(dss[jb["scene"]], boxes[jb["scene"]], jb["r"], jb["ng"], jb["nc"], cam, jb["frames"], jb["seed"],
                           k_global=64, r_global=jb["r"], k_caustic=64, r_caustic=jb["r"], max_paths=3 * jb["ng"],
                           gather_flags=pm355.KNN_TUNED if tuned else 0)
                nxt += 1
            if srv.step() == 0 and nxt < len(jobs):
                time.sleep(max(0.0, min(0.01, jobs[nxt]["at"] - (time.perf_counter() - t0))))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = srv.latencies() * 1e3
        info = srv.info()
        out.setdefault(f"{rate}_jobs_per_s", {})[cname] = {
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "frames": int(info["frames"]), "jobs": len(jobs), "queries_per_s": info["queries"] / wall,
            "wall_s": wall}
        srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--photons", type=int, default=16_000_000)
    ap.add_argument("--queries", type=int, default=4_000_000)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--radius", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--sort-queries", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--frames", type=int, default=12, help="per-frame Cornell-box frames for p50/p99 (0 = skip)")
    ap.add_argument("--batch-jobs", type=int, default=256, help="jobs of the batched gather mix (0 = skip)")
    ap.add_argument("--ref-frames", type=int, default=5,
                    help="frames of the reference application's CornellBox-Sphere frame (0 = skip)")
    ap.add_argument("--ref-frame-probe", action="store_true",
                    help="only render the reference frame (3 frames per grid cell size, no CPU part) and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB on the device 3 times (the byte-counter scale reference) and exit")
    ap.add_argument("--cell-over-r", type=float, default=0.0, help="grid cell size / r (0 = autotune)")
    ap.add_argument("--variant", default="", help="gather variant (empty = autotune)")
    ap.add_argument("--trace-probe", action="store_true",
                    help="only trace the Cornell-box photon maps twice (for rocprofv3 counter runs) and exit")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="after tuning, launch the tuned gather once on the cell-ordered block and exit")
    ap.add_argument("--serve-jobs", type=int, default=50, help="render jobs per arrival rate for the served form (0 = skip)")
    ap.add_argument("--frame-scenes", type=int, default=256, help="scenes of the batched frames mix (0 = skip)")
    args = ap.parse_args()

    build = None
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the photon path "
              f"runs on a single MI355X; the photon path shards by query, tile and frame over a replicated photon map)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_probe:
        a = torch.ones(256 * 1024 * 1024, device="cuda")  # 1 GiB of float32
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}), flush=True)
        return

    import pm355
    from pm355.workloads import gather_microbench

    if args.trace_probe:
        # two C2 photon traces (2M global + 500k caustic, Cornell box with
        # BVHs) and nothing else, for rocprofv3 counter runs
        from pm355.scenes import cornell_box

        ctx = pm355.Context(0)
        ds = ctx.upload_scene(cornell_box(), accel=True)
        for seed in (1, 2):
            g, c, st = ctx.trace_photons(ds, seed, 2_000_000, 500_000, 10, max_paths=200_000_000)
        torch.cuda.synchronize()
        print(json.dumps({"trace_probe": {"paths": int(st[0]), "global": int(st[1]), "caustic": int(st[2])}}),
              flush=True)
        ctx.close()
        return

    if args.ref_frame_probe:
        ctx = pm355.Context(0)
        print(json.dumps({"reference_frame": reference_frame_bench(ctx, 3)}), flush=True)
        ctx.close()
        return

    log("generating workload")
    ph, q, qn, aabb = gather_microbench(seed=args.seed, n_photons=args.photons, n_queries=args.queries)
    r, k = args.radius, args.k
    ctx = pm355.Context(0)
    dev = ctx.device
    d_ph = torch.from_numpy(ph).to(dev)
    d_q = torch.from_numpy(q).to(dev)
    d_n = torch.from_numpy(qn).to(dev)
    nq = q.shape[1]
    stream = torch.cuda.current_stream(dev)

    def timed(fn, reps):
        # one untimed call first: it grows scratch and allocates outputs
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # Autotune (the library's pm_tune_gather): the grid's cell size relative
    # to r and the gather variant, jointly, for the step's time: each cell
    # size's grid rebuild and each variant's fused gather (queries sorted
    # inside the call), timed with HIP events on the launch stream.
    if args.cell_over_r or args.variant:
        frac = args.cell_over_r or 1.0
        vname = args.variant or "stream512_chunked"
        tune, tuning = {}, None
    else:
        tuning = ctx.tune_gather(d_ph, aabb, d_q, d_n, r, k)
        tune = ctx.tuning_table()
        frac = tuning["cell_over_r"]
        vname = {v: n_ for n_, v in pm355.KNN_VARIANTS.items()}[tuning["flags"]]
        best_ms = tuning["ms"]
        # Timings within 2% are noise: then keep the configuration whose
        # counters were profiled (profiles/), so the bench line can report them.
        tpath0 = os.path.join(ROOT, "profiles", "gather_microbench_16M_4M_k64_counters.json")
        if os.path.exists(tpath0):
            t0 = json.load(open(tpath0))
            pf, pv = t0.get("cell_over_r"), t0.get("variant")
            pms = tune.get(f"{pf:.3f}", {}).get(pv) if pf is not None else None
            if pms is not None and pms <= best_ms * 1.02:
                frac, vname = pf, pv
    cell = r * frac
    vflags = pm355.KNN_VARIANTS[vname]
    flags = (pm355.KNN_SORT_QUERIES if args.sort_queries else 0) | vflags
    log(f"cell/r {frac:.3f}, variant {vname}: {tune}")
    g = ctx.grid_build(d_ph, aabb, cell_size=cell)
    order = ctx.query_order(g, d_q).long()
    d_qs = d_q[:, order].contiguous()
    d_ns = d_n[:, order].contiguous()
    torch.cuda.synchronize()
    if args.traffic_probe:
        # one launch of the tuned gather on the cell-ordered block, for
        # rocprofv3 --pmc runs (FETCH_SIZE / WRITE_SIZE per dispatch)
        ctx.gather_radiance(g, d_qs, d_ns, r, k, flags=vflags)
        torch.cuda.synchronize()
        print(json.dumps({"traffic_probe": {"cell_over_r": frac, "variant": vname}}), flush=True)
        return

    # The step's grid is resident: each step rebuilds it in place from the
    # photon block (pm_grid_rebuild), as a per-frame renderer keeps its grid
    # memory; a fresh build's allocation is timed separately below.
    g_step = ctx.grid_build(d_ph, aabb, cell_size=cell)

    def step():
        g_step.rebuild(d_ph)
        return ctx.gather_radiance(g_step, d_q, d_n, r, k, flags=flags, want_count=True)

    log("warmup")
    for _ in range(args.warmup):
        rgb, cnt = step()
    torch.cuda.synchronize()

    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rgb, cnt = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g_step.close()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    log(f"{ms_per_step:.3f} ms/step")

    # Component timing with HIP events on the stream the kernels run on.
    reps = max(3, args.steps)
    fresh_build_ms = timed(lambda: ctx.grid_build(d_ph, aabb, cell_size=cell).close(), reps)
    build_ms = timed(lambda: g.rebuild(d_ph), reps)
    order_ms = timed(lambda: ctx.query_order(g, d_q), reps)
    gather_ms = timed(lambda: ctx.gather_radiance(g, d_qs, d_ns, r, k, flags=vflags), reps)

    # Algorithmic bytes of one fused gather launch, per query:
    # 12 (query) + 8*C_q (start + size) + 12*P_q
    # (candidate positions) + 8*kept (index + distance) + 24*kept (power +
    # direction of the kept photons) + 12 (normal) + 12 (RGB out), with C_q
    # and P_q the cells and photons of the reference's candidate box at the
    # reference cell size s = r (exact counts, as the oracle's counting pass
    # gives them). The kernel itself walks that box pruned to what can reach
    # r at the tuned cell size ("walked" below), and most of its reads hit in
    # L2 / MALL, so these bytes per second can exceed the HBM peak: they
    # are reported beside the roofline, not as it.
    rgb, cnt = ctx.gather_radiance(g, d_qs, d_ns, r, k, flags=vflags, want_count=True)
    torch.cuda.synchronize()
    kept = int(cnt.to(torch.int64).sum().item())
    g_ref = ctx.grid_build(d_ph, aabb, cell_size=r)
    cells_ref, cands_ref = ctx.box_stats(g_ref, d_q, r, reference=True)
    g_ref.close()
    alg_bytes = nq * (12 + 12 + 12) + 8 * cells_ref + 12 * cands_ref + (8 + 24) * kept
    cells, cands = ctx.box_stats(g, d_q, r)
    walked_bytes = nq * (12 + 12 + 12) + 8 * cells + 12 * cands + 24 * kept

    # Bytes per launch at each memory level, from the committed rocprofv3
    # counter summary of this configuration (pm355.profiling; flagged stale
    # when the HIP sources changed since it was measured): L2 read bytes
    # (TCC_READ_SECTORS x 32) and HBM bytes (FETCH_SIZE / WRITE_SIZE scaled
    # by a 1 GiB copy). The candidate positions of neighbouring cell-ordered
    # queries are read from L2 / the Infinity Cache, so L2 is the level that
    # serves the gather's bytes and the roofline is drawn there. Without a
    # matching summary, the live walked-box bytes (every candidate position,
    # cell bound and kept record the kernel reads) stand in for the L2 bytes.
    from pm355 import profiling

    cpath = os.path.join(ROOT, "profiles", "gather_microbench_16M_4M_k64_counters.json")
    cs = profiling.load(cpath, variant=vname, cell_over_r=float(frac))
    t_s = gather_ms * 1e-3
    if cs and "l2_read_bytes_per_dispatch" in cs:
        l2_bytes, l2_src = cs["l2_read_bytes_per_dispatch"], "TCC_READ_SECTORS x 32 B (profiles/" + \
            os.path.basename(cpath) + (", STALE: measured on other HIP sources)" if cs["stale"] else ")")
    else:
        l2_bytes, l2_src = walked_bytes, "walked-box bytes (no counter summary for this configuration)"
    traffic = None
    if cs and "hbm_read_bytes_per_dispatch" in cs and "hbm_write_bytes_per_dispatch" in cs:
        traffic = cs["hbm_read_bytes_per_dispatch"] + cs["hbm_write_bytes_per_dispatch"]
    achieved = l2_bytes / t_s / 1e9

    value = nq / (ms_per_step * 1e-3)
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "gather microbenchmark: 16M photons (70% cube faces, 30% 64 floor clusters sigma=0.005), "
                        f"4M face queries, k=64, r=0.01, grid cell={frac:.3f}r (autotuned); step = grid rebuild in "
                        "place + fused kNN gather/radiance (queries cell-sorted inside the call)",
            "photons": args.photons,
            "queries": nq,
            "k": k,
            "radius": r,
            "cell_over_r": frac,
            "sort_queries": bool(args.sort_queries),
            "gather_variant": vname,
        },
        "roofline": {
            "bound": "l2",
            "achieved": achieved,
            "peak": profiling.L2_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / profiling.L2_PEAK_GBS,
            "traffic": traffic,
            "kernel": {"stream512_chunked": "k_knn_stream512", "stream512": "k_knn_stream512",
                       "inner": "k_knn_stream512", "stream1024": "k_knn_stream1024",
                       "two_pass": "k_knn_two_pass"}.get(vname, vname)
                      + f" (variant {vname}, cell {frac:.3f}r; fused gather + radiance, cell-ordered queries)",
            "kernel_ms": gather_ms,
            "l2_read_bytes": l2_bytes,
            "l2_read_bytes_source": l2_src,
            "peak_source": "aggregate L2 read bandwidth: 8 XCDs x 16 channels x 128 B/clk x 2.4 GHz",
            "hbm_GBs": traffic / t_s / 1e9 if traffic else None,
            "hbm_frac": traffic / t_s / 1e9 / HBM_PEAK_GBS if traffic else None,
            "l2_hit_rate": cs.get("l2_hit_rate") if cs else None,
            "l1_hit_rate": cs.get("l1_hit_rate") if cs else None,
            "l2_read_latency_cycles": cs.get("l2_read_latency_cycles") if cs else None,
            "counters_stale": cs["stale"] if cs else None,
            # the profiled launch's instructions per query (one wave per
            # query) and the VALU lane utilisation (active lanes per VALU
            # instruction / 64)
            "insts_per_query": insts_per_query(cs),
            "valu_lane_util": valu_lane_util(cs),
            "walked_bytes": walked_bytes,
            "alg_bytes": alg_bytes,
            "alg_bytes_definition": "sum over queries of 12 + 8 C_q + 12 P_q + 32 kept + 24 (reference box at "
                                    "cell = r); its rate exceeds the HBM peak because neighbouring queries share "
                                    "candidates in L2 / MALL",
            "alg_GBs": alg_bytes / t_s / 1e9,
        },
        "components_ms": {
            "grid_build": build_ms,
            "grid_build_fresh_allocation": fresh_build_ms,
            "query_order": order_ms,
            "gather_radiance": gather_ms,
        },
        "autotune_ms": tune,
        "autotune": None if tuning is None else {
            "source": "pm_tune_gather (library): grid rebuild + sorted fused gather per cell size and variant",
            "chosen": {"cell_over_r": tuning["cell_over_r"], "variant": vname, "ms": tuning["ms"]},
            "default": {"cell_over_r": 1.0, "variant": "stream512_chunked", "ms": tuning["ms_default"]},
            "speedup_vs_default": tuning["ms_default"] / tuning["ms"]},
        "gather_only_queries_per_s": nq / (gather_ms * 1e-3),
        "box_stats": {"reference_box_cell_r": {"cells": cells_ref, "candidates": cands_ref},
                      "walked_tuned_cell": {"cells": cells, "candidates": cands,
                                            "bytes": walked_bytes,
                                            "GBs": walked_bytes / (gather_ms * 1e-3) / 1e9},
                      "kept": kept},
    }
    if frac != 1.0:
        # the reference-compatible cell = r beside the tuned cell size
        gr = ctx.grid_build(d_ph, aabb, cell_size=r)
        orr = ctx.query_order(gr, d_q).long()
        qs, ns = d_q[:, orr].contiguous(), d_n[:, orr].contiguous()
        ms_r = timed(lambda: ctx.gather_radiance(gr, qs, ns, r, k, flags=vflags), 3)
        b_r = timed(lambda: gr.rebuild(d_ph), 3)
        res["cell_equals_r"] = {"gather_ms": ms_r, "grid_build_ms": b_r, "step_queries_per_s_estimate":
                                nq / ((ms_r + b_r + order_ms) * 1e-3)}
        gr.close()
        del qs, ns

    # kd-tree build over the same 16M photons (PhotonMap), HIP events, and
    # the kd-tree's own kNN query (PhotonMap::locate_q) over the same 4M
    # queries in cell order: neighbour lists (index + distance), k, r as above
    kd_ms = timed(lambda: ctx.kdtree_build(d_ph).close(), 3)
    res["components_ms"]["kdtree_build_16M"] = kd_ms
    kt = ctx.kdtree_build(d_ph)
    kd_knn_ms = timed(lambda: ctx.kdtree_knn(kt, d_qs, r, k), 2)
    grid_knn_ms = timed(lambda: ctx.knn(g, d_qs, r, k, flags=vflags), 2)
    kt.close()
    res["components_ms"]["kdtree_knn_4M"] = kd_knn_ms
    res["components_ms"]["grid_knn_4M"] = grid_knn_ms
    res["kdtree_knn_queries_per_s"] = nq / (kd_knn_ms * 1e-3)
    res["grid_knn_queries_per_s"] = nq / (grid_knn_ms * 1e-3)

    # Measured copy peak (a STREAM-like device copy of 1 GiB: bytes read +
    # written per second) beside the published 8 TB/s, and the build
    # rooflines. Algorithmic bytes: grid build N*(12 + 36 + 36) + 12*G
    # (position read for the key; record read + SoA write; starts, sizes and
    # the cell-bounds pass per cell); kd-tree build N*L*12 + N*72 with
    # L = ceil(log2(N + 1)) levels (split key, index read and write per
    # level; the final record permutation).
    a = torch.empty(256 * 1024 * 1024, device=dev)
    b = torch.empty_like(a)
    copy_ms = timed(lambda: b.copy_(a), 5)
    copy_gbs = 2 * a.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del a, b
    n_ph = d_ph.shape[1]
    n_cells = g.info.n_cells
    grid_bytes = n_ph * (12 + 36 + 36) + 12 * n_cells
    levels = int(np.ceil(np.log2(n_ph + 1)))
    kd_bytes = n_ph * levels * 12 + n_ph * 72
    res["copy_peak_GBs"] = copy_gbs
    res["build_roofline"] = {
        "grid_build": {"ms": build_ms, "alg_bytes": grid_bytes, "cells": n_cells,
                       "achieved_GBs": grid_bytes / (build_ms * 1e-3) / 1e9,
                       "frac": grid_bytes / (build_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kdtree_build_16M": {"ms": kd_ms, "alg_bytes": kd_bytes, "levels": levels,
                             "achieved_GBs": kd_bytes / (kd_ms * 1e-3) / 1e9,
                             "frac": kd_bytes / (kd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    g.close()
    if args.frames > 0:
        log(f"per-frame form: {args.frames} Cornell-box frames")
        res["per_frame"] = frame_bench(ctx, args.frames)
    if not args.no_cpu_baseline:
        build = native_oracle()
    if args.ref_frames > 0:
        log(f"reference frame: {args.ref_frames} CornellBox-Sphere frames")
        res["reference_frame"] = reference_frame_bench(ctx, args.ref_frames,
                                                       0 if args.no_cpu_baseline else host_threads())
    if args.batch_jobs > 0:
        log(f"batched form: {args.batch_jobs}-job gather mix")
        res["batched"] = batch_bench(ctx, args.batch_jobs)
    if args.frame_scenes > 0:
        log(f"batched frames: {args.frame_scenes}-scene frame mix")
        res["batched_frames"] = frames_batch_bench(ctx, args.frame_scenes)
        ctx.trim()  # the mix's grid and frame blocks
    if args.serve_jobs > 0:
        log(f"served form: {args.serve_jobs} jobs per arrival rate")
        res["served"] = {"workload": "Poisson arrivals of render jobs: 6 resident scenes of the batched distribution "
                                     "(unit box + 1-8 spheres of 200-20k triangles, BVHs), 50k-8M global photons + 5% "
                                     "caustic re-traced per frame (at most 3 paths per global photon), 128^2..1920x1080, "
                                     "1-30 frames, k=64, r U[0.01,0.05]; per rate: jobs as given (cell = r) and tuned "
                                     "(tunings of the batched mix and the headline)",
                         **serve_bench(ctx, n_jobs=args.serve_jobs)}
    if not args.no_cpu_baseline:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(ph, q, qn, aabb, r, k, build)
        res["cpu_stages"] = cpu_stages(res["cpu_baseline"]["cores"], build)
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
