"""CPU: OBJ/MTL loader (pm355.objload) and the headless image writers.

The shipped CornellBox-Sphere files (the reference's Models/cornell_box_sphere,
public domain) are checked against facts read off the files by other means:
the face, vertex and group counts of the OBJ text, the scene bounds, the
libc strtof value of every number, and the materials the MTL defines on top
of the Material() defaults (Material.cuh:22-30). Small synthetic files pin
the reference loader's rules one at a time (ObjLoader.cpp:4-306).
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN
from pm355.image import canvas_rows, read_ppm, write_pfm, write_ppm
from pm355.objload import load_material, load_obj
from pm355.scenes import reference_camera

SCENE_DIR = os.path.join(GOLDEN, "cornell_box_sphere")
OBJ = os.path.join(SCENE_DIR, "CornellBox-Sphere.obj")
MTL = os.path.join(SCENE_DIR, "CornellBox-Sphere.mtl")
f32 = np.float32


@pytest.fixture(scope="module")
def shipped():
    return load_obj(SCENE_DIR, "CornellBox-Sphere.obj")


def _obj_lines():
    with open(OBJ) as f:
        return f.read().split("\n")


def test_shipped_models_and_counts(shipped):
    s = shipped
    assert s.names == ["leftSphere", "rightSphere", "floor", "ceiling", "backWall", "rightWall", "leftWall",
                       "light"]
    assert s.light_model == 7
    lines = _obj_lines()
    n_faces = sum(1 for ln in lines if ln.startswith("f "))
    assert s.tris9.shape == (n_faces, 9) == (2188, 9)
    # per-group face counts from the text: faces between one `g` line and the next
    counts, cur = [], None
    for ln in lines:
        if ln.startswith("g "):
            if cur is not None:
                counts.append(cur)
            cur = 0
        elif ln.startswith("f ") and cur is not None:
            cur += 1
    counts.append(cur)
    assert list(s.models["tri_end"] - s.models["tri_begin"]) == counts
    assert list(s.models["tri_begin"][1:]) == list(s.models["tri_end"][:-1])
    assert list(s.models["id"]) == list(range(8))


def test_shipped_bounds_and_vertices(shipped):
    s = shipped
    # bounds of every `v` line (all are used by faces)
    v = np.array([[float(t) for t in ln.split()[1:4]] for ln in _obj_lines() if ln.startswith("v ")], np.float32)
    assert np.array_equal(np.array(s.aabb(), np.float32), np.concatenate([v.min(0), v.max(0)]))
    assert s.aabb() == pytest.approx((-1.02, 0.0, -1.04, 1.0, 1.59, 0.99), abs=1e-6)
    # the light: 2 triangles on the quad of the last 4 vertices, normals from one vn
    m = s.models[s.light_model]
    lt = s.tris9[m["tri_begin"]:m["tri_end"]].reshape(-1, 3, 3)
    assert np.array_equal(np.unique(lt.reshape(-1, 3), axis=0), np.unique(v[-4:], axis=0))
    ln_ = s.normals9[m["tri_begin"]:m["tri_end"]].reshape(-1, 3)
    assert np.all(ln_ == ln_[0])
    # every triangle's vertex normal faces its own side: unit length as in the file
    nn = np.linalg.norm(s.normals9.reshape(-1, 3).astype(np.float64), axis=1)
    assert np.all(np.abs(nn - 1) < 2e-3)


def test_shipped_numbers_parse_like_strtof():
    libc = ctypes.CDLL(None)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    toks = set()
    for path in (OBJ, MTL):
        with open(path) as f:
            for ln in f:
                t = ln.split()
                if t and t[0] in ("v", "vn", "vt", "Ka", "Kd", "Ks", "Ke", "Ns", "Ni", "d", "Tr"):
                    toks.update(t[1:4] if t[0] not in ("Ns", "Ni", "d", "Tr") else t[1:2])
    toks.discard("#")
    assert len(toks) > 500
    for t in toks:
        assert np.float32(float(t)).view(np.uint32) == np.float32(libc.strtof(t.encode(), None)).view(np.uint32), t


def test_shipped_materials(shipped):
    m = {n: shipped.models[i] for i, n in enumerate(shipped.names)}
    # mirror sphere: Kd, Ks, Ns given; opaque and IOR stay at the defaults
    assert np.array_equal(m["leftSphere"]["kd"], f32([0.01] * 3))
    assert np.array_equal(m["leftSphere"]["ks"], f32([0.95] * 3))
    assert m["leftSphere"]["opaque"] == f32(1) and m["leftSphere"]["refr_index"] == f32(1)
    assert m["leftSphere"]["shininess"] == f32(0)
    # glass sphere: Tr 0.9 -> opaque = 1 - 0.9 in float32, Ni 1.5
    assert m["rightSphere"]["opaque"] == f32(1) - f32(0.9)
    assert m["rightSphere"]["refr_index"] == f32(1.5)
    # walls: Ks 0 overrides the default specular 1; IOR only where Ni is given
    assert np.array_equal(m["backWall"]["kd"], f32([1, 1, 0]))
    for n in ("floor", "ceiling", "backWall", "rightWall", "leftWall", "light"):
        assert np.array_equal(m[n]["ks"], f32([0, 0, 0])), n
        assert m[n]["opaque"] == f32(1), n
    assert m["floor"]["refr_index"] == f32(1.5) and m["rightWall"]["refr_index"] == f32(1.0)
    assert np.array_equal(m["rightWall"]["kd"], f32([0.161, 0.133, 0.427]))
    # the light emits Ke 10, nothing else does
    assert np.array_equal(m["light"]["ke"], f32([10, 10, 10]))
    assert all(not np.any(m[n]["ke"]) for n in shipped.names if n != "light")


def test_material_defaults_and_first_block(tmp_path):
    p = tmp_path / "m.mtl"
    p.write_text("newmtl a\nKd 0.5 0.25 0.125\nnewmtl a\nKd 9 9 9\nnewmtl b\nTr 0.25\nd 0.5\n")
    a = load_material(str(p), "a")
    assert np.array_equal(a["diffuse"], f32([0.5, 0.25, 0.125]))  # first block wins
    assert np.array_equal(a["specular"], f32([1, 1, 1]))  # Material() default
    assert a["shininess"] == f32(16) and a["opaque"] == f32(1) and a["refr_index"] == f32(1)
    b = load_material(str(p), "b")
    assert b["opaque"] == f32(0.5)  # the later of Tr / d wins
    z = load_material(str(p), "missing")
    assert np.array_equal(z["diffuse"], f32([1, 1, 1]))


def _write(tmp_path, obj, mtl="newmtl m1\nKd 0.5 0.5 0.5\nKs 0 0 0\nnewmtl m2\nKd 0.1 0.2 0.3\n"):
    (tmp_path / "s.mtl").write_text(mtl)
    (tmp_path / "s.obj").write_text(obj)
    return str(tmp_path / "s.obj")


_QUAD = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"


def test_quad_split_negative_indices_and_names(tmp_path):
    obj = ("mtllib s.mtl\n" + _QUAD +
           "g a\nusemtl m1\nf 1//1 2//1 3//1 4//1\n"
           "g a\nf -4//-1 -3//-1 -2//-1\n"        # same name -> a1; material carried over
           "o light\nusemtl m2\nf 1/1/1 2/1/1 3/1/1\n")
    s = load_obj(_write(tmp_path, obj))
    assert s.names == ["a", "a1", "light"] and s.light_model == 2
    t = s.tris9.reshape(-1, 3, 3)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    assert np.array_equal(t[0], v[[0, 1, 3]]) and np.array_equal(t[1], v[[1, 2, 3]])  # Model.cuh:318-340
    assert np.array_equal(t[2], v[[0, 1, 2]])
    assert list(s.models["tri_end"]) == [2, 3, 4]
    assert np.array_equal(s.models[1]["kd"], f32([0.5] * 3))  # carried over
    assert np.array_equal(s.models[2]["kd"], f32([0.1, 0.2, 0.3]))
    assert np.array_equal(s.models[2]["ks"], f32([1, 1, 1]))  # usemtl resets to defaults


def test_running_name_keeps_suffix(tmp_path):
    obj = "mtllib s.mtl\n" + _QUAD + "g x\nf 1//1 2//1 3//1\n#\nf 1//1 2//1 3//1\n#\nf 1//1 2//1 3//1\ng light\nf 1//1 2//1 3//1\n"
    s = load_obj(_write(tmp_path, obj))
    # the second model renames x -> x1; the third finds x1 taken -> x11
    assert s.names == ["x", "x1", "x11", "light"]


def test_whitespace_line_repeats_keyword(tmp_path):
    # "  " after a `v` line pushes the last vertex again (the stream keeps the
    # keyword and fails the reads), shifting later indices by one
    obj = "mtllib s.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\n  \nv 5 5 5\nvn 0 0 1\ng light\nf 1//1 2//1 5//1\n"
    s = load_obj(_write(tmp_path, obj))
    assert np.array_equal(s.tris9[0, 6:9], f32([5, 5, 5]))
    obj2 = "mtllib s.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\n\nv 5 5 5\nvn 0 0 1\ng light\nf 1//1 2//1 4//1\n"
    s2 = load_obj(_write(tmp_path, obj2))
    assert np.array_equal(s2.tris9[0, 6:9], f32([5, 5, 5]))  # an empty line clears the keyword


def test_faces_without_normals_get_geometric_normal(tmp_path):
    obj = "mtllib s.mtl\nv 0 0 0\nv 2 0 0\nv 0 0 2\ng light\nf 1 2 3\n"
    s = load_obj(_write(tmp_path, obj))
    assert np.array_equal(s.normals9[0].reshape(3, 3), np.tile(f32([0, -1, 0]), (3, 1)))


@pytest.mark.parametrize("obj,msg", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nv 2 2 0\ng light\nf 1 2 3 4 5\n", "vertices"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\ng light\nf 0 1 2\n", "index 0"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\ng light\nf 1 2 9\n", "out of range"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\ng lamp\nf 1 2 3\n", "light"),
    ("v 0 0 0\n", "no faces"),
])
def test_rejected_inputs(tmp_path, obj, msg):
    with pytest.raises(ValueError, match=msg):
        load_obj(_write(tmp_path, obj))


def test_reference_camera_matches_camera_conventions():
    c = reference_camera()
    assert (int(c["width"]), int(c["height"])) == (1024, 768)
    assert np.array_equal(c["forward"], f32([0, 0, -1]))
    assert np.array_equal(c["right"], f32([1, 0, 0])) and np.array_equal(c["up"], f32([0, 1, 0]))
    # 60 degree vertical field of view: viewport height 2 tan(30 deg)
    assert float(c["viewport_height"]) == pytest.approx(2 * np.tan(np.pi / 6), rel=1e-6)
    assert float(c["viewport_width"]) == pytest.approx(float(c["viewport_height"]) * 4 / 3, rel=1e-6)


def test_ppm_and_pfm_orientation(tmp_path):
    W, H = 5, 3
    canvas = np.zeros((H, W, 3), np.uint8)
    canvas[0, :, 0] = 255          # canvas row 0 = bottom of the picture
    canvas[H - 1, 0] = (1, 2, 3)   # top-left pixel
    p = str(tmp_path / "a.ppm")
    write_ppm(p, canvas.reshape(-1), W, H)
    img = read_ppm(p)
    assert img.shape == (H, W, 3)
    assert np.all(img[H - 1, :, 0] == 255) and tuple(img[0, 0]) == (1, 2, 3)
    assert np.array_equal(img, canvas_rows(canvas.reshape(-1), W, H))
    with open(p, "rb") as f:
        assert f.read(11) == b"P6\n5 3\n255\n"
    planes = np.arange(3 * W * H, dtype=np.float32).reshape(3, W * H)
    q = str(tmp_path / "a.pfm")
    write_pfm(q, planes, W, H)
    with open(q, "rb") as f:
        data = f.read()
    hdr = b"PF\n5 3\n-1.0\n"
    assert data.startswith(hdr)
    px = np.frombuffer(data[len(hdr):], "<f4").reshape(H, W, 3)
    assert np.array_equal(px[0, 0], planes[:, 0]) and np.array_equal(px[2, 4], planes[:, 14])


@pytest.fixture(scope="module")
def cpp_dump(tmp_path_factory):
    """tests/objload_dump.cpp (the C++ loader of host/scene_io.hpp) built
    with g++ for the host."""
    import subprocess

    from conftest import ROOT

    d = tmp_path_factory.mktemp("cppobj")
    exe = str(d / "objload_dump")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "objload_dump.cpp")], check=True)
    return exe


def _cpp_load(exe, obj, out, W=1024, H=768):
    import subprocess

    from pm355.scenes import PM_CAMERA_DTYPE, PM_MODEL_DTYPE

    os.makedirs(out, exist_ok=True)
    r = subprocess.run([exe, obj, out, str(W), str(H)], capture_output=True, text=True)
    if r.returncode != 0:
        return r.stderr
    rd = lambda n, dt: np.fromfile(os.path.join(out, n), dtype=dt)
    return dict(tris9=rd("tris9.f32", np.float32).reshape(-1, 9), normals9=rd("normals9.f32", np.float32).reshape(-1, 9),
                models=rd("models.bin", PM_MODEL_DTYPE), light=int(rd("light.i32", np.int32)[0]),
                names=open(os.path.join(out, "names.txt")).read().split("\n")[:-1],
                camera=rd("camera.bin", PM_CAMERA_DTYPE)[0])


def _same(c, s):
    assert c["names"] == s.names and c["light"] == s.light_model
    assert np.array_equal(c["tris9"].view(np.uint32), s.tris9.view(np.uint32))
    assert np.array_equal(c["normals9"].view(np.uint32), s.normals9.view(np.uint32))
    assert c["models"].tobytes() == s.models.tobytes()


def test_cpp_loader_equals_python_on_shipped_scene(cpp_dump, shipped, tmp_path):
    c = _cpp_load(cpp_dump, OBJ, str(tmp_path))
    _same(c, shipped)
    for W, H in ((1024, 768), (96, 72)):
        c = _cpp_load(cpp_dump, OBJ, str(tmp_path / f"c{W}"), W, H)
        assert c["camera"].tobytes() == np.ascontiguousarray(reference_camera(W, H)).tobytes()


@pytest.mark.parametrize("obj", [
    "mtllib s.mtl\n" + _QUAD + "g a\nusemtl m1\nf 1//1 2//1 3//1 4//1\ng a\nf -4//-1 -3//-1 -2//-1\n"
    "o light\nusemtl m2\nf 1/1/1 2/1/1 3/1/1\n",
    "mtllib s.mtl\n" + _QUAD + "g x\nf 1//1 2//1 3//1\n#\nf 1//1 2//1 3//1\n#\nf 1//1 2//1 3//1\ng light\nf 1//1 2//1 3//1\n",
    "mtllib s.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\n  \nv 5 5 5\nvn 0 0 1\ng light\nf 1//1 2//1 5//1\n",
    "mtllib s.mtl\nv 0 0 0\nv 2 0 0\nv 0 0 2\ng light\nusemtl nosuch\nf 1 2 3\nf 1 3 2\n",
    "mtllib s.mtl\r\nv 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nvn 0 0 1\r\n\r\ng light\r\nusemtl m1\r\nf 1//1 2//1 3//1\r\n",
])
def test_cpp_loader_equals_python_rules(cpp_dump, tmp_path, obj):
    path = _write(tmp_path, obj)
    _same(_cpp_load(cpp_dump, path, str(tmp_path / "o")), load_obj(path))


@pytest.mark.parametrize("obj", ["v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nv 2 2 0\ng light\nf 1 2 3 4 5\n",
                                 "v 0 0 0\nv 1 0 0\nv 0 1 0\ng light\nf 0 1 2\n",
                                 "v 0 0 0\nv 1 0 0\nv 0 1 0\ng lamp\nf 1 2 3\n"])
def test_cpp_loader_rejects_like_python(cpp_dump, tmp_path, obj):
    path = _write(tmp_path, obj)
    assert isinstance(_cpp_load(cpp_dump, path, str(tmp_path / "o")), str)
    with pytest.raises(ValueError):
        load_obj(path)


def _random_obj(rng, path):
    """A random OBJ/MTL pair and the scene it must load as: groups of
    triangles and quads in the v//vn, v/vt/vn and negative-index forms,
    vertices declared between groups, materials with every keyword."""
    n_mat = int(rng.integers(2, 5))
    mats = []
    mtl = []
    for i in range(n_mat):
        m = dict(diffuse=rng.random(3).astype(np.float32), specular=rng.random(3).astype(np.float32),
                 emission=np.zeros(3, np.float32), shininess=np.float32(rng.integers(0, 64)),
                 refr_index=np.float32(1.0), opaque=np.float32(1.0))
        lines = [f"newmtl mat{i}", "Kd %r %r %r" % tuple(float(x) for x in m["diffuse"]),
                 "Ks %r %r %r" % tuple(float(x) for x in m["specular"]), f"Ns {int(m['shininess'])}"]
        if rng.random() < 0.5:
            m["refr_index"] = np.float32(1.5)
            lines.append("Ni 1.5")
        if rng.random() < 0.5:
            tr = np.float32(rng.integers(1, 9) / 10)
            m["opaque"] = np.float32(np.float32(1.0) - tr)
            lines.append(f"Tr {float(tr)!r}")
        mats.append(m)
        mtl += lines + [""]
    light = dict(diffuse=np.full(3, 0.5, np.float32), specular=np.ones(3, np.float32),
                 emission=np.full(3, 4.0, np.float32), shininess=np.float32(16), refr_index=np.float32(1),
                 opaque=np.float32(1))
    mtl += ["newmtl lamp", "Kd 0.5 0.5 0.5", "Ke 4 4 4"]
    obj = ["mtllib r.mtl"]
    pos, nrm = [], []
    tris, nrms, models, names = [], [], [], []
    n_groups = int(rng.integers(2, 6))
    for gi in range(n_groups + 1):
        is_light = gi == n_groups
        nv = int(rng.integers(4, 12))
        base = len(pos)
        for _ in range(nv):
            v = (rng.random(3) * 4 - 2).astype(np.float32)
            pos.append(v)
            obj.append("v %r %r %r" % tuple(float(x) for x in v))
        nb = len(nrm)
        nn = rng.standard_normal((2, 3))
        for v in nn:
            v = (v / np.linalg.norm(v)).astype(np.float32)
            nrm.append(v)
            obj.append("vn %r %r %r" % tuple(float(x) for x in v))
        name = "light" if is_light else f"grp{gi}"
        obj.append(f"g {name}")
        mi = int(rng.integers(0, n_mat))
        obj.append("usemtl lamp" if is_light else f"usemtl mat{mi}")
        t0 = len(tris)
        for _ in range(int(rng.integers(1, 6))):
            quad = rng.random() < 0.4
            idx = rng.choice(nv, 4 if quad else 3, replace=False) + base
            ni = [int(rng.integers(0, 2)) + nb for _ in idx]
            form = int(rng.integers(0, 3))
            toks = []
            for p_i, n_i in zip(idx, ni):
                if form == 0:
                    toks.append(f"{p_i + 1}//{n_i + 1}")
                elif form == 1:
                    toks.append(f"{p_i + 1}/1/{n_i + 1}")
                else:  # negative: counted back from the vertices read so far
                    toks.append(f"{p_i - len(pos)}//{n_i - len(nrm)}")
            obj.append("f " + " ".join(toks))
            order = [(0, 1, 3), (1, 2, 3)] if quad else [(0, 1, 2)]
            for a, b, c in order:
                tris.append(np.concatenate([pos[idx[a]], pos[idx[b]], pos[idx[c]]]))
                nrms.append(np.concatenate([nrm[ni[a]], nrm[ni[b]], nrm[ni[c]]]))
        m = light if is_light else mats[mi]
        models.append((t0, len(tris), m))
        names.append(name)
        obj.append("")
    with open(os.path.join(path, "r.mtl"), "w") as f:
        f.write("\n".join(mtl) + "\n")
    with open(os.path.join(path, "r.obj"), "w") as f:
        f.write("\n".join(obj) + "\n")
    return np.array(tris, np.float32), np.array(nrms, np.float32), models, names


@pytest.mark.parametrize("seed", range(6))
def test_generated_obj_loads_as_built(cpp_dump, tmp_path, seed):
    tris, nrms, models, names = _random_obj(np.random.default_rng(seed), str(tmp_path))
    s = load_obj(str(tmp_path / "r.obj"))
    assert s.names == names and s.light_model == len(names) - 1
    assert np.array_equal(s.tris9, tris) and np.array_equal(s.normals9, nrms)
    for rec, (t0, t1, m) in zip(s.models, models):
        assert (rec["tri_begin"], rec["tri_end"]) == (t0, t1)
        assert np.array_equal(rec["kd"], m["diffuse"]) and np.array_equal(rec["ks"], m["specular"])
        assert np.array_equal(rec["ke"], m["emission"])
        assert rec["shininess"] == m["shininess"] and rec["refr_index"] == m["refr_index"]
        assert rec["opaque"] == m["opaque"]
    _same(_cpp_load(cpp_dump, str(tmp_path / "r.obj"), str(tmp_path / "o")), s)
