"""The C++ façade (cuda-photon-mapping_amd/host/pm355.hpp) end to end: the
demo program builds a PhotonGrid, queries find_nearests, builds and queries a
PhotonMap through the reference's names; its outputs must equal the oracle's
on the same photons."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
from conftest import PKG

pytestmark = pytest.mark.gpu


def test_facade_demo_matches_oracle(tmp_path):
    exe = os.path.join(PKG, "facade_demo")
    n, nq = 60000, 3000
    r = subprocess.run([exe, str(tmp_path), str(n), str(nq), "5"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rd = lambda name, dt: np.fromfile(os.path.join(tmp_path, name), dtype=dt)
    ph = rd("photons.f32", np.float32).reshape(9, n)
    q = rd("queries.f32", np.float32).reshape(3, nq)
    og = O.OracleGrid(ph, (0, 0, 0, 1, 1, 1), cell_size=0.03)
    assert np.array_equal(rd("grid_starts.u32", np.uint32), og.starts)
    assert np.array_equal(rd("grid_sizes.u32", np.uint32), og.counts)
    oi, od, oc = og.knn(q, 0.05, 32)
    assert np.array_equal(rd("knn_count.u32", np.uint32), oc)
    assert np.array_equal(rd("knn_idx.u32", np.uint32).reshape(nq, 32), oi)
    assert np.array_equal(rd("knn_dist.f32", np.float32).reshape(nq, 32).view(np.uint32), od.view(np.uint32))
    ki, kp = O.kdtree_build(ph[0:3])
    assert np.array_equal(rd("kd_idx.u32", np.uint32), ki)
    assert np.array_equal(rd("kd_plane.i8", np.int8), kp)
    qi, qd, qc = O.kdtree_knn(ph[0:3], ki, kp, q, 0.05, 32)
    assert np.array_equal(rd("kdknn_count.u32", np.uint32), qc)
    assert np.array_equal(rd("kdknn_idx.u32", np.uint32).reshape(nq, 32), qi)
    assert np.array_equal(rd("kdknn_dist.f32", np.float32).reshape(nq, 32).view(np.uint32), qd.view(np.uint32))


def test_facade_render_obj_equals_python_path(ctx, tmp_path):
    """`facade_demo render`: the C++ loader, DeviceScene, trace_photons, two
    PhotonGrids and render_frame on the shipped CornellBox-Sphere give the
    photon maps and canvas of the Python path (pm355.objload + Context),
    which the shipped-scene tests check against the oracle."""
    from conftest import GOLDEN
    from pm355.image import read_ppm
    from pm355.objload import load_obj
    from pm355.scenes import reference_camera

    exe = os.path.join(PKG, "facade_demo")
    obj = os.path.join(GOLDEN, "cornell_box_sphere", "CornellBox-Sphere.obj")
    W, H, ng, nc, seed = 128, 96, 100_000, 10_000, 3
    r = subprocess.run([exe, "render", obj, str(tmp_path), str(W), str(H), str(ng), str(nc), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    scene = load_obj(obj)
    ds = ctx.upload_scene(scene, accel=True)
    g, c, st = ctx.trace_photons(ds, seed, ng, nc, 10)
    aabb = scene.aabb()
    gg, gc = ctx.grid_build(g, aabb), ctx.grid_build(c, aabb)
    _, rgb8, _ = ctx.render_frame(ds, gg, gc, reference_camera(W, H), 4, 2000, 0.1, 400, 0.2)
    ctx.sync()
    rd = lambda name, dt: np.fromfile(os.path.join(tmp_path, name), dtype=dt)
    assert np.array_equal(rd("global9.f32", np.uint32), g.cpu().numpy().reshape(-1).view(np.uint32))
    assert np.array_equal(rd("caustic9.f32", np.uint32), c.cpu().numpy().reshape(-1).view(np.uint32))
    assert np.array_equal(rd("rgb8.u8", np.uint8), rgb8.cpu().numpy())
    # the maps persisted in the reference's photon-map file format
    from pm355.workloads import read_photon_file

    assert np.array_equal(read_photon_file(os.path.join(tmp_path, "global_map.bin")).view(np.uint32),
                          g.cpu().numpy().view(np.uint32))
    assert np.array_equal(read_photon_file(os.path.join(tmp_path, "caustic_map.bin")).view(np.uint32),
                          c.cpu().numpy().view(np.uint32))
    assert np.array_equal(read_ppm(os.path.join(tmp_path, "frame.ppm"))[::-1].reshape(-1), rgb8.cpu().numpy())
    gg.close()
    gc.close()


def test_photon_files_through_the_c_abi(ctx, tmp_path):
    """pm_photon_file_read / _write against the independent numpy reader and
    writer: the reference fixture global_map_test.txt (100 photons) loads
    as numpy reads it; a 2.5M-photon block (three streaming chunks) round
    trips bit for bit; errors for a missing file, a truncated file and a
    capacity below the count."""
    import ctypes as C

    import torch
    import pm355
    from conftest import GOLDEN
    from pm355.workloads import read_photon_file, write_photon_file

    fx = os.path.join(GOLDEN, "global_map_test.bin")
    got = ctx.photon_file_read(fx)
    assert np.array_equal(got.cpu().numpy().view(np.uint32), read_photon_file(fx).view(np.uint32))
    ph = torch.rand((9, 2_500_000), device="cuda")
    p = os.path.join(tmp_path, "big.bin")
    ctx.photon_file_write(p, ph)
    assert os.path.getsize(p) == 8 + 36 * ph.shape[1]
    assert np.array_equal(read_photon_file(p).view(np.uint32), ph.cpu().numpy().view(np.uint32))
    back = ctx.photon_file_read(p)
    assert torch.equal(back, ph)
    q = os.path.join(tmp_path, "np.bin")
    write_photon_file(q, ph[:, :1000].cpu().numpy())
    assert torch.equal(ctx.photon_file_read(q), ph[:, :1000])
    with pytest.raises(pm355.PMError):
        ctx.photon_file_read(os.path.join(tmp_path, "missing.bin"))
    with open(os.path.join(tmp_path, "short.bin"), "wb") as f:
        f.write(open(p, "rb").read()[: 8 + 36 * 10 + 5])
    with pytest.raises(pm355.PMError):
        ctx.photon_file_read(os.path.join(tmp_path, "short.bin"))
    out = torch.empty((9, 10), device="cuda")
    n = C.c_uint32()
    rc = ctx.lib.pm_photon_file_read(ctx.h, os.fsencode(p), out.data_ptr(), 10, C.byref(n))
    assert rc == pm355.PM_ERR_CAPACITY and n.value == 0
