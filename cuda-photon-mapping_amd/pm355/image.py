"""Headless image output for rendered frames.

The reference shows each frame by copying its uchar3 canvas into an OpenGL
pixel buffer (Raytracer.cuh:95-107, 277, 292; Window.cuh), with canvas row
y = 0 at the bottom of the window (Camera::generate_ray_direction maps
pixel_y = 0 to the lower viewport edge, Camera.cuh:67-81). These writers
store the same canvas as a binary PPM (P6) or as the HDR-mapped float planes
in a PFM (Pf/PF), both top row first.
"""
from __future__ import annotations

import numpy as np


def _host(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def canvas_rows(rgb8, width: int, height: int) -> np.ndarray:
    """uint8 [H, W, 3] image, top row first, from a canvas of W*H uchar3 with
    pixel (x, y) at y*W + x and y = 0 at the bottom."""
    c = _host(rgb8).astype(np.uint8, copy=False).reshape(height, width, 3)
    return c[::-1]


def write_ppm(path: str, rgb8, width: int, height: int) -> None:
    """Binary PPM (P6, maxval 255) of a uchar3 canvas."""
    img = np.ascontiguousarray(canvas_rows(rgb8, width, height))
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (width, height))
        f.write(img.tobytes())


def read_ppm(path: str) -> np.ndarray:
    """uint8 [H, W, 3] from a binary PPM written by write_ppm (top row first)."""
    with open(path, "rb") as f:
        data = f.read()
    toks, pos = [], 0
    while len(toks) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            pos = data.index(b"\n", pos) + 1
            continue
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        toks.append(data[pos:end])
        pos = end
    if toks[0] != b"P6" or int(toks[3]) != 255:
        raise ValueError(f"{path}: not a P6 PPM with maxval 255")
    w, h = int(toks[1]), int(toks[2])
    pos += 1  # the single whitespace byte before the raster
    return np.frombuffer(data, np.uint8, w * h * 3, pos).reshape(h, w, 3)


def write_pfm(path: str, rgb3, width: int, height: int) -> None:
    """Colour PFM ("PF", little-endian float32) of the 3 HDR planes
    [3, W*H]. PFM stores the bottom row first, which is canvas row 0."""
    p = _host(rgb3).astype(np.float32, copy=False).reshape(3, height, width)
    img = np.ascontiguousarray(np.moveaxis(p, 0, -1)).astype("<f4")
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (width, height))
        f.write(img.tobytes())
