"""Headless renderer: OBJ/MTL scene -> photon maps -> frame -> PPM.

The command-line counterpart of the reference application's main loop
(kernel.cu:8-40, Initializer.cu:179-201, Drawer.cu:18-101) without the
window: load the scene, trace the global and caustic photon maps, build both
grids, render one or more frames and write each canvas as a binary PPM (and
optionally the HDR-mapped floats as PFM). Defaults are the reference's
(REFERENCE_FRAME, reference_camera); `--cell-over-r` picks the grid cell
size relative to each gather radius instead of the reference's sizing rule
(neighbour sets and images do not depend on it).

    python -m pm355.render_obj tests/golden/cornell_box_sphere/CornellBox-Sphere.obj frame.ppm
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def main(argv=None) -> int:
    from .scenes import REFERENCE_FRAME as RF

    ap = argparse.ArgumentParser(prog="python -m pm355.render_obj", description=__doc__.splitlines()[0])
    ap.add_argument("obj", help="OBJ file (its MTL is found beside it)")
    ap.add_argument("out", help="output PPM; with --frames N > 1, '_<f>' is inserted before the extension")
    ap.add_argument("--pfm", action="store_true", help="also write the HDR-mapped floats as .pfm")
    ap.add_argument("--width", type=int, default=RF["width"])
    ap.add_argument("--height", type=int, default=RF["height"])
    ap.add_argument("--global-photons", type=int, default=RF["n_global"])
    ap.add_argument("--caustic-photons", type=int, default=RF["n_caustic"])
    ap.add_argument("--max-bounces", type=int, default=RF["max_bounces"])
    ap.add_argument("--max-depth", type=int, default=RF["max_depth"])
    ap.add_argument("--k-global", type=int, default=RF["k_global"])
    ap.add_argument("--r-global", type=float, default=RF["r_global"])
    ap.add_argument("--k-caustic", type=int, default=RF["k_caustic"])
    ap.add_argument("--r-caustic", type=float, default=RF["r_caustic"])
    ap.add_argument("--cell-over-r", type=float, default=0.0,
                    help="grid cell size / gather radius (0 = the reference's sizing rule)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--frames", type=int, default=1, help="frames, each with freshly traced photon maps")
    ap.add_argument("--no-bvh", action="store_true", help="brute-force intersection (as the reference)")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)

    import torch

    from . import Context
    from .image import write_pfm, write_ppm
    from .objload import load_obj
    from .scenes import camera

    scene = load_obj(a.obj)
    ctx = Context(a.device)
    ds = ctx.upload_scene(scene, accel=not a.no_bvh)
    aabb = scene.aabb()
    # the reference's camera placement, on any canvas size
    cam = camera((-0.00999999046, 0.795000017, 2.35000001), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 60.0,
                 a.width, a.height)
    root, ext = os.path.splitext(a.out)
    for f in range(a.frames):
        t0 = time.perf_counter()
        g, c, st = ctx.trace_photons(ds, a.seed + f, a.global_photons, a.caustic_photons, a.max_bounces)
        gg = ctx.grid_build(g, aabb, cell_size=a.cell_over_r * a.r_global)
        gc = ctx.grid_build(c, aabb, cell_size=a.cell_over_r * a.r_caustic) if st[2] > 0 else None
        rgb, rgb8, hits = ctx.render_frame(ds, gg, gc, cam, a.max_depth, a.k_global, a.r_global, a.k_caustic,
                                           a.r_caustic)
        torch.cuda.synchronize(ctx.device)
        ms = (time.perf_counter() - t0) * 1e3
        path = a.out if a.frames == 1 else f"{root}_{f}{ext}"
        write_ppm(path, rgb8, a.width, a.height)
        if a.pfm:
            write_pfm(os.path.splitext(path)[0] + ".pfm", rgb, a.width, a.height)
        print(json.dumps({"frame": f, "file": path, "ms": round(ms, 3), "paths": st[0], "global": st[1],
                          "caustic": st[2], "eye_ray_hits": hits}), flush=True)
        gg.close()
        if gc is not None:
            gc.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
