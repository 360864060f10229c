"""Wavefront OBJ/MTL scene loader producing a pm355 Scene.

Follows the reference's loader line by line in what it keeps
(ObjLoader.cpp:123-306, materials ObjLoader.cpp:4-83):

* a model is the run of consecutive `f` lines; the first non-`f` line after
  it (comment, blank line, `v`, `g`, `usemtl` ...) closes it
  (ObjLoader.cpp:169-186), and so does the end of the file (:289-300);
* the model is named by the last `g` or `o` line (default "n1"); a name
  already taken gets one '1' appended, and the suffix stays on the running
  name (:174-177);
* `usemtl` resets the material to Material() defaults (ambient 0.2, diffuse
  1, specular 1, emission 0, shininess 16, IOR 1, opaque 1;
  Material.cuh:4-30) and then applies the first `newmtl` block of that name:
  Kd, Ks, Ke, Ns, Ni; `d` sets opaque and `Tr` sets opaque = 1 - Tr
  (ObjLoader.cpp:4-83). A model keeps the material of the one before it when
  no `usemtl` comes between them (:179-181);
* position and normal indices are 1-based, negative ones count back from the
  vertices read so far (:96-112);
* a line holding only whitespace leaves the previous line's keyword in
  force with no operands read, so the keyword's action repeats with the last
  values parsed (the std::stringstream behaviour of the reference loop); a
  line of length zero clears the keyword. Lines end at '\n' only, as with
  std::getline (a '\r' before it is whitespace).

The model emitting the photons is the one named exactly "light"
(Initializer.cu:62-84).

Where the reference has no working behaviour this loader chooses one and
says so:
* faces written `f v v v` or `f v/vt ...`: the reference's index counter
  assigns the second and third numbers of such faces to texture and normal
  indices (ObjLoader.cpp:234-262); here they are positions, as in the OBJ
  format;
* faces without normal indices: the reference leaves their normals
  uninitialised (:106-112); here they get the face's geometric normal;
* quads: the reference's Quad models are split into (v0, v1, v3) and
  (v1, v2, v3) for intersection (Model.cuh:318-340); here every 4-vertex face
  becomes those two triangles, so a model may mix triangles and quads
  (the reference types a whole model by its last face, :273-281);
* a model with no faces (a bare `f` line) takes its name but is dropped;
* faces of more than 4 vertices (the reference's FlatPolygon, which has no
  primitive size, ModelConstructInfo.cu:6-12) and index 0 raise ValueError.

Texture coordinates, `s` and `map_Kd` are read past: nothing on the photon
path uses them. Numbers are parsed to float32 round-to-nearest (what
`std::istream >> float` gives).
"""
from __future__ import annotations

import os

import numpy as np

from .scenes import PM_MODEL_DTYPE, Scene, _model

# Material() defaults (Material.cuh:22-30)
_MATERIAL_DEFAULT = dict(ambient=(0.2, 0.2, 0.2), diffuse=(1.0, 1.0, 1.0), specular=(1.0, 1.0, 1.0),
                         emission=(0.0, 0.0, 0.0), shininess=16.0, refr_index=1.0, opaque=1.0)


def _f32(tok: str) -> np.float32:
    return np.float32(float(tok))


def _vec3(ops, stale):
    """Three floats from the operands, or the stale value on a failed read."""
    if len(ops) < 3:
        if ops:
            raise ValueError(f"malformed vector operands {ops!r}")
        return stale
    return np.array([_f32(ops[0]), _f32(ops[1]), _f32(ops[2])], np.float32)


def _split(line: str, prefix: str):
    """(prefix, operands) of one line; see the module note on blank lines."""
    if len(line) == 0:
        return "", []
    toks = line.split()
    if not toks:
        return prefix, []
    return toks[0], toks[1:]


def load_material(mtl_path: str, name: str) -> dict:
    """Material `name` from an MTL file over Material() defaults
    (ObjLoader.cpp:4-83)."""
    mat = {k: (np.array(v, np.float32) if isinstance(v, tuple) else np.float32(v))
           for k, v in _MATERIAL_DEFAULT.items()}
    with open(mtl_path, "r", newline="") as f:
        lines = f.read().split("\n")
    prefix = ""
    tv = np.zeros(3, np.float32)
    tf = np.float32(0)
    processing = False
    for line in lines:
        prefix, ops = _split(line, prefix)
        if prefix == "#":
            continue
        if prefix == "newmtl":
            if processing:
                break
            if (ops[0] if ops else "") != name:
                continue
            processing = True
        elif not processing:
            continue
        elif prefix in ("Ka", "Kd", "Ks", "Ke"):
            tv = _vec3(ops, tv)
            key = {"Ka": "ambient", "Kd": "diffuse", "Ks": "specular", "Ke": "emission"}[prefix]
            mat[key] = tv.copy()
        elif prefix in ("d", "Tr", "Ns", "Ni"):
            if ops:
                tf = _f32(ops[0])
            if prefix == "d":
                mat["opaque"] = tf
            elif prefix == "Tr":
                mat["opaque"] = np.float32(np.float32(1.0) - tf)
            elif prefix == "Ns":
                mat["shininess"] = tf
            else:
                mat["refr_index"] = tf
    return mat


def _index(tok: str, n: int, what: str) -> int:
    i = int(tok)
    if i == 0:
        raise ValueError(f"{what} index 0 (OBJ indices are 1-based)")
    j = i - 1 if i > 0 else n + i
    if not 0 <= j < n:
        raise ValueError(f"{what} index {i} out of range ({n} defined)")
    return j


def load_obj(path: str, fname: str | None = None) -> Scene:
    """Scene from an OBJ file and the MTL it names (loadOBJ,
    ObjLoader.cpp:123-306). `path` is the file, or its folder when `fname`
    is given (the reference's (path, fname) pair)."""
    if fname is not None:
        path = os.path.join(path, fname)
    folder = os.path.dirname(path)
    with open(path, "r", newline="") as f:
        lines = f.read().split("\n")

    pos, nrm = [], []
    tv = np.zeros(3, np.float32)
    mtl_fname = None
    cur_name = "n1"
    names_seen: set = set()
    names: list = []
    cur_mat = None          # material dict of the model being read
    cur_faces: list = []    # [(pos idx tuple, nrm idx tuple or None)] of the open model
    finished = False
    models = []             # (name, faces, material)
    prefix = ""
    mat_cache: dict = {}

    def material(name):
        key = (mtl_fname, name)
        if key not in mat_cache:
            if mtl_fname is None:
                raise ValueError("usemtl before mtllib")
            mat_cache[key] = load_material(os.path.join(folder, mtl_fname), name)
        return mat_cache[key]

    def close_model():
        nonlocal cur_name, cur_faces
        if cur_name in names_seen:
            cur_name += "1"
        names_seen.add(cur_name)
        if cur_faces:  # a bare `f` line makes an empty model: named, not kept
            names.append(cur_name)
            models.append((cur_name, cur_faces, cur_mat))
        cur_faces = []

    for line in lines:
        prefix, ops = _split(line, prefix)
        if finished and prefix != "f":
            close_model()
            finished = False
        if prefix == "mtllib":
            if ops:
                mtl_fname = ops[0]
        elif prefix in ("g", "o"):
            if ops:
                cur_name = ops[0]
        elif prefix == "usemtl":
            cur_mat = material(ops[0] if ops else "")
        elif prefix == "v":
            tv = _vec3(ops, tv)
            pos.append(tv.copy())
        elif prefix == "vn":
            tv = _vec3(ops, tv)
            nrm.append(tv.copy())
        elif prefix == "f":
            finished = True
            if not ops:
                continue
            if len(ops) > 4:
                raise ValueError(f"face with {len(ops)} vertices: only triangles and quads are supported")
            if len(ops) < 3:
                raise ValueError(f"face with {len(ops)} vertices")
            pi, ni = [], []
            for t in ops:
                parts = t.split("/")
                pi.append(_index(parts[0], len(pos), "position"))
                ni.append(_index(parts[2], len(nrm), "normal") if len(parts) > 2 and parts[2] else None)
            if len(ops) == 3:
                cur_faces.append((pi, ni))
            else:  # Model.cuh:318-340 split
                for a, b, c in ((0, 1, 3), (1, 2, 3)):
                    cur_faces.append(([pi[a], pi[b], pi[c]], [ni[a], ni[b], ni[c]]))
    if finished:
        close_model()
    if not models:
        raise ValueError(f"{path}: no faces")

    P = np.array(pos, np.float32).reshape(-1, 3)
    N = np.array(nrm, np.float32).reshape(-1, 3)
    tris, nrms, recs = [], [], []
    t0 = 0
    for mi, (name, faces, mat) in enumerate(models):
        if mat is None:
            mat = {k: (np.array(v, np.float32) if isinstance(v, tuple) else np.float32(v))
                   for k, v in _MATERIAL_DEFAULT.items()}
        for pi, ni in faces:
            v = P[pi]
            tris.append(v.reshape(9))
            if any(i is None for i in ni):
                g = np.cross((v[1] - v[0]).astype(np.float64), (v[2] - v[0]).astype(np.float64))
                ln = np.linalg.norm(g)
                g = (g / ln if ln > 0 else g).astype(np.float32)
                nrms.append(np.array([N[i] if i is not None else g for i in ni], np.float32).reshape(9))
            else:
                nrms.append(N[ni].reshape(9))
        recs.append(_model(t0, t0 + len(faces), kd=tuple(mat["diffuse"]), ks=tuple(mat["specular"]),
                           ke=tuple(mat["emission"]), ior=float(mat["refr_index"]), opaque=float(mat["opaque"]),
                           shininess=float(mat["shininess"]), mid=mi))
        t0 += len(faces)
    if "light" not in names:
        raise ValueError(f"{path}: no model named 'light' (the photon source)")
    return Scene(np.array(tris, np.float32).reshape(-1, 9), np.array(nrms, np.float32).reshape(-1, 9),
                 np.array(recs, PM_MODEL_DTYPE), light_model=names.index("light"), names=names)
