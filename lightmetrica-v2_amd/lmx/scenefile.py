"""Scene files of the reference (the YAML schema of main.cpp:539-767 and
scene3.cpp:98-432; assets as assets.cpp:53-104 resolves them) loaded into
the renderer scene of `lmx.render.Scene`, and rendered as main.cpp's
`render` command does (renderer from the `renderer` node, film saved with
the film asset's type).

Semantics kept from the reference:
* node transforms (scene3.cpp:159-294): `matrix` is 16 numbers filling
  m[i/4][i%4] -- column-major text, so the translation is numbers 13-15;
  `lookat` gives the columns vx, vy, vz, eye with vz = normalize(eye -
  center), vx = normalize(cross(up, vz)), vy = cross(vz, vx); `translate`,
  `rotate` (axis, angle in degrees) and `scale` compose as T R S; a node's
  world transform is parent x local; the normal transform is the upper 3x3
  of transpose(inverse(M));
* mesh vertices are M (p, 1) without a division by w (accel_bvh_sahbin.cpp:
  98-100); triangles are enumerated primitive by primitive in traversal
  order, faces in mesh order, and area lights in that order;
* the pinhole sensor's position is M (0, 0, 0, 1) and its axes the first
  three columns of M (sensor_pinhole.cpp:47-54), fov in degrees (default 45),
  aspect from its film asset;
* a node without `bsdf` has the null BSDF; `light` and `sensor` exclude each
  other.

Decisions where the reference is undefined or inconsistent:
* a node without `transform` has the identity, and so do its children as
  their parent transform (the reference passes an uninitialised matrix);
* OBJ shapes are appended in file order (TriangleMesh_Obj prepends each
  shape's vertices while offsetting its indices as if appended,
  trianglemesh_obj.cpp:83-90, which breaks files of several shapes);
* matrices and vectors need exactly 16 / 3 numbers.
"""
import math
import os

import numpy as np

from . import meshfile
from .scenes import MAT_DIFFUSE, MAT_GGX, MAT_NULL

VERSION = "1.1.0"                          # the only version main.cpp:562-597 accepts


class SceneFileError(ValueError):
    pass


def _nums(s, n, what):
    v = [float(t) for t in str(s).split()]
    if len(v) != n:
        raise SceneFileError(f"{what}: expected {n} numbers, got {len(v)}")
    return np.array(v, np.float64)


def _mat4_text(s):
    """16 numbers filling m[i/4][i%4]: column i/4, row i%4 (property.h:192-200)."""
    v = _nums(s, 16, "matrix")
    return v.reshape(4, 4).T.copy()          # M[row, col]


def _translate(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m


def _rotate(deg, axis):
    """Math::Rotate (math.h:2293-2320) as a row/column matrix."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    ang = deg * math.pi / 180.0
    c, s = math.cos(ang), math.sin(ang)
    t = (1.0 - c) * a
    rot = np.eye(4)
    # the reference fills rot[col][row]
    rc = np.array([[c + t[0] * a[0], t[0] * a[1] + s * a[2], t[0] * a[2] - s * a[1]],
                   [t[1] * a[0] - s * a[2], c + t[1] * a[1], t[1] * a[2] + s * a[0]],
                   [t[2] * a[0] + s * a[1], t[2] * a[1] - s * a[0], c + t[2] * a[2]]])
    rot[:3, :3] = rc.T
    return rot


def _scale(v):
    m = np.eye(4)
    m[0, 0], m[1, 1], m[2, 2] = v
    return m


def parse_transform(node):
    """The local transform of a `transform` node (scene3.cpp:175-279)."""
    if node is None:
        return np.eye(4)
    if node.get("matrix") is not None:
        return _mat4_text(node["matrix"])
    if node.get("lookat") is not None:
        la = node["lookat"]
        for k in ("eye", "center", "up"):
            if la.get(k) is None:
                raise SceneFileError(f"lookat: missing '{k}'")
        eye, center, up = (_nums(la[k], 3, "lookat " + k) for k in ("eye", "center", "up"))
        vz = eye - center
        vz /= np.linalg.norm(vz)
        vx = np.cross(up, vz)
        vx /= np.linalg.norm(vx)
        vy = np.cross(vz, vx)
        m = np.eye(4)
        m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = vx, vy, vz, eye
        return m
    t, r, s = node.get("translate"), node.get("rotate"), node.get("scale")
    m = np.eye(4)
    if t is not None:
        m = m @ _translate(_nums(t, 3, "translate"))
    if r is not None:
        if r.get("angle") is None or r.get("axis") is None:
            raise SceneFileError("rotate: needs 'axis' and 'angle'")
        m = m @ _rotate(float(r["angle"]), _nums(r["axis"], 3, "rotate axis"))
    if s is not None:
        m = m @ _scale(_nums(s, 3, "scale"))
    return m


def parse_nodes(scene_node):
    """Primitives of a scene3 params node in traversal order: dicts with id,
    transform (4x4 world), mesh, bsdf, light, sensor (asset ids or None)."""
    if not isinstance(scene_node, dict) or scene_node.get("nodes") is None:
        raise SceneFileError("Missing 'nodes' node")
    prims = []

    def walk(n, parent):
        n = n or {}
        if n.get("transform") is None:
            world = np.eye(4)
        else:
            world = parent @ parse_transform(n["transform"])
        if n.get("light") is not None and n.get("sensor") is not None:
            raise SceneFileError("'light' and 'sensor' node cannot be used in the same time")
        prims.append({"id": None if n.get("id") is None else str(n["id"]), "transform": world,
                      "mesh": n.get("mesh"), "bsdf": n.get("bsdf"), "light": n.get("light"),
                      "sensor": n.get("sensor")})
        for c in n.get("child") or []:
            walk(c, world)

    for n in scene_node["nodes"]:
        walk(n, np.eye(4))
    sid = scene_node.get("sensor")
    if sid is None:
        raise SceneFileError("Missing 'sensor' node")
    ids = {p["id"]: i for i, p in enumerate(prims) if p["id"] is not None}
    if str(sid) not in ids:
        raise SceneFileError(f"Missing primitive ID: {sid}")
    return prims, ids[str(sid)]


# ---- OBJ ------------------------------------------------------------------

def _fix(idx, n):
    return idx - 1 if idx > 0 else (0 if idx == 0 else n + idx)


def load_obj(path):
    """Triangle mesh of an OBJ file as the reference's vendored tinyobjloader
    reads it: shapes end at `g`, `o` and `usemtl` lines; within a shape every
    distinct (v, vt, vn) triple becomes one vertex; polygons are fanned
    (0, k-1, k); indices are 1-based, negative ones relative.  Returns the
    dict of meshfile.load_raw_mesh."""
    v, vn, vt = [], [], []
    shapes, group = [], []

    def flush():
        if not group:
            return
        cache, P, N, T, F = {}, [], [], [], []

        def vert(t):
            if t not in cache:
                P.append(v[t[0]])
                if t[2] >= 0:
                    N.append(vn[t[2]])
                if t[1] >= 0:
                    T.append(vt[t[1]])
                cache[t] = len(P) - 1
            return cache[t]
        for face in group:
            for k in range(2, len(face)):
                F.append((vert(face[0]), vert(face[k - 1]), vert(face[k])))
        shapes.append((P, N, T, F))
        group.clear()

    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                v.append([float(x) for x in tok[1:4]])
            elif tok[0] == "vn":
                vn.append([float(x) for x in tok[1:4]])
            elif tok[0] == "vt":
                vt.append([float(x) for x in tok[1:3]])
            elif tok[0] == "f":
                face = []
                for t in tok[1:]:
                    parts = t.split("/")
                    vi = _fix(int(parts[0]), len(v))
                    ti = _fix(int(parts[1]), len(vt)) if len(parts) > 1 and parts[1] else -1
                    ni = _fix(int(parts[2]), len(vn)) if len(parts) > 2 and parts[2] else -1
                    face.append((vi, ti, ni))
                group.append(face)
            elif tok[0] in ("g", "o", "usemtl"):
                flush()
    flush()
    if not shapes:
        raise SceneFileError(f"{path}: no faces")
    has_n, has_t = bool(shapes[0][1]), bool(shapes[0][2])
    if any(bool(s[1]) != has_n or bool(s[2]) != has_t for s in shapes):
        raise SceneFileError("Inconsistency of normal or texcoords")
    ps, ns, ts, fs = [], [], [], []
    for P, N, T, F in shapes:
        base = len(ps)
        ps += P; ns += N; ts += T
        fs += [(a + base, b + base, c + base) for a, b, c in F]
    return {"positions": np.array(ps, np.float32).reshape(-1, 3),
            "normals": np.array(ns, np.float32).reshape(-1, 3) if has_n else None,
            "texcoords": np.array(ts, np.float32).reshape(-1, 2) if has_t else None,
            "faces": np.array(fs, np.uint32).reshape(-1, 3)}


# ---- assets and the renderer scene ------------------------------------------

def _asset(assets, aid, interface):
    a = assets.get(str(aid))
    if a is None:
        raise SceneFileError(f"Missing asset '{aid}'")
    if str(a.get("interface", "")).lower() != interface:
        raise SceneFileError(f"Asset '{aid}' is not a {interface}")
    return a


def _vec3(p, key, default):
    return tuple(_nums(p[key], 3, key)) if p.get(key) is not None else default


def load_scene(source, base_dir=None):
    """Reads a scene file (a path, or YAML text with `base_dir` for relative
    mesh paths).  Returns dict(scene=<renderer scene dict>, renderer=<type,
    params>, film=<w, h, type>, primitives=<parse_nodes output>)."""
    import yaml
    if os.path.exists(str(source)):
        base_dir = base_dir or os.path.dirname(os.path.abspath(str(source)))
        with open(source) as fh:
            doc = yaml.safe_load(fh)
    else:
        doc = yaml.safe_load(source)
    base_dir = base_dir or "."
    root = (doc or {}).get("lightmetrica")
    if root is None:
        raise SceneFileError("Missing 'lightmetrica' node")
    if str(root.get("version", "")) != VERSION:
        raise SceneFileError(f"Invalid version '{root.get('version')}': this loader reads {VERSION}")
    assets = (root.get("assets") or {}).get("params") or {}
    scene_node = (root.get("scene") or {}).get("params")
    prims, sensor_idx = parse_nodes(scene_node)
    verts, tris, norms, tri_mat, mats, lights = [], [], [], [], [], []
    any_normals = False
    meshes = {}

    def mesh_of(mid):
        if mid not in meshes:
            a = _asset(assets, mid, "trianglemesh")
            p = a.get("params") or {}
            if a.get("type") == "raw":
                meshes[mid] = meshfile.load_raw_mesh(p)
            elif a.get("type") == "obj":
                meshes[mid] = load_obj(os.path.join(base_dir, str(p["path"])))
            else:
                raise SceneFileError(f"trianglemesh::{a.get('type')} is not supported")
        return meshes[mid]

    for p in prims:
        if p["mesh"] is not None and mesh_of(str(p["mesh"]))["normals"] is not None:
            any_normals = True
    nvert = ntri = 0
    for p in prims:
        if p["mesh"] is None:
            continue
        m = mesh_of(str(p["mesh"]))
        M = p["transform"]
        P = np.concatenate([m["positions"].astype(np.float64), np.ones((m["positions"].shape[0], 1))], axis=1)
        W = (P @ M.T)[:, :3].astype(np.float32)                      # M (p, 1), w dropped
        F = m["faces"].astype(np.int64)
        if any_normals and m["normals"] is None:
            # geometric normals: one vertex per triangle corner
            W = W[F.reshape(-1)]
            g = np.cross(W[1::3] - W[0::3], W[2::3] - W[0::3])
            N = np.repeat(g / np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-30), 3, axis=0)
            F = np.arange(W.shape[0]).reshape(-1, 3)
        elif any_normals:
            nt = np.linalg.inv(M).T[:3, :3]
            N = (m["normals"].astype(np.float64) @ nt.T).astype(np.float32)
        else:
            N = None
        # material: the bsdf, and the light's Le
        mat = dict(type=MAT_NULL, R=(0.0, 0.0, 0.0), exponent=0.0, Le=(0.0, 0.0, 0.0))
        if p["bsdf"] is not None:
            b = _asset(assets, p["bsdf"], "bsdf")
            bp = b.get("params") or {}
            if bp.get("TexR") is not None:
                raise SceneFileError("textures (TexR) are not supported")
            if b.get("type") == "diffuse":
                mat.update(type=MAT_DIFFUSE, R=_vec3(bp, "R", (0.0, 0.0, 0.0)))
            elif b.get("type") == "cook_torrance":
                mat.update(type=MAT_GGX, R=_vec3(bp, "R", (0.0, 0.0, 0.0)),
                           eta=_vec3(bp, "eta", (0.14, 0.129, 0.1585)), k=_vec3(bp, "k", (4.58625, 3.348125, 2.329375)),
                           roughness=float(bp.get("roughness", 0.1)))
            else:
                raise SceneFileError(f"bsdf::{b.get('type')} is not supported")
        if p["light"] is not None:
            la = _asset(assets, p["light"], "light")
            if la.get("type") != "area":
                raise SceneFileError(f"light::{la.get('type')} is not supported")
            mat["Le"] = _vec3(la.get("params") or {}, "Le", (0.0, 0.0, 0.0))
            lights.append((ntri, F.shape[0]))
        key = repr(sorted(mat.items()))
        idx = next((i for i, q in enumerate(mats) if repr(sorted(q.items())) == key), None)
        if idx is None:
            mats.append(mat)
            idx = len(mats) - 1
        tris.append((F + nvert).astype(np.uint32))
        verts.append(W)
        if N is not None:
            norms.append(N.astype(np.float32))
        tri_mat.append(np.full(F.shape[0], idx, np.uint32))
        nvert += W.shape[0]
        ntri += F.shape[0]
    sp = prims[sensor_idx]
    if sp["sensor"] is None:
        raise SceneFileError("the main sensor node has no 'sensor' asset")
    sa = _asset(assets, sp["sensor"], "sensor")
    if sa.get("type") != "pinhole":
        raise SceneFileError(f"sensor::{sa.get('type')} is not supported")
    spar = sa.get("params") or {}
    film = _asset(assets, spar.get("film"), "film")
    fp = film.get("params") or {}
    M = sp["transform"]
    camera = dict(frame=(M[:3, 3].astype(np.float32), M[:3, 0].astype(np.float32), M[:3, 1].astype(np.float32),
                         M[:3, 2].astype(np.float32)),
                  fov_deg=float(spar.get("fov", 45.0)), width=int(fp["w"]), height=int(fp["h"]))
    if not tris:
        raise SceneFileError("the scene has no triangles")
    scene = dict(verts=np.concatenate(verts), tris=np.concatenate(tris), normals=np.concatenate(norms) if any_normals else None,
                 tri_mat=np.concatenate(tri_mat), materials=mats, lights=lights, camera=camera)
    rnode = root.get("renderer") or {}
    return dict(scene=scene, renderer={"type": rnode.get("type"), "params": rnode.get("params") or {}},
                film={"w": int(fp["w"]), "h": int(fp["h"]), "type": str(fp.get("type", "radiancehdr"))},
                primitives=prims)


def render(loaded, seed=42, raster_mode=None):
    """Renders a load_scene result with its renderer node (pt, ptdirect,
    ptmis, bdpt, pm, ppm, sppm, ltdirect, raycast; vcm in its bdpt, vcm
    and bdpm modes) and the reference's parameter defaults.
    Returns (film tensor, stats)."""
    from . import render as R
    rs = R.Scene(loaded["scene"])
    t, p = loaded["renderer"]["type"], loaded["renderer"]["params"]
    n = int(p.get("num_samples", 10000000))
    rm = R.RASTER_RANDOM if raster_mode is None else raster_mode
    if t in ("pt", "ptdirect", "ptmis"):
        fn = {"pt": rs.render_pt, "ptdirect": rs.render_ptdirect, "ptmis": rs.render_ptmis}[t]
        film, st, _ = fn(n, max_num_vertices=int(p.get("max_num_vertices", -1)),
                         min_num_vertices=int(p.get("min_num_vertices", 0)), seed=seed, raster_mode=rm)
    elif t == "bdpt":
        mis = R.MIS_SIMPLE if str(p.get("mis", "powerheuristics")) == "simple" else R.MIS_POWER
        film, st, _ = rs.render_bdpt(n, max_num_vertices=int(p.get("max_num_vertices", -1)),
                                     min_num_vertices=int(p.get("min_num_vertices", 0)), seed=seed, raster_mode=rm,
                                     mis=mis)
    elif t == "vcm":
        # renderer_vcm.cpp:558-577: every mode runs num_iteration_pass passes
        # of num_eye_trace_samples eye samples and keeps their mean image
        mode = str(p.get("mode", "vcm"))
        passes = int(p.get("num_iteration_pass", 100))
        ne = int(p.get("num_eye_trace_samples", 10000))
        common = dict(max_num_vertices=int(p.get("max_num_vertices", 10)),
                      min_num_vertices=int(p.get("min_num_vertices", 0)), seed=seed, raster_mode=rm)
        if mode == "bdpt":
            # independent samples: one render of passes x eye samples is the passes' mean
            film, st, _ = rs.render_vcm_bdpt(passes * ne, **common)
        elif mode in ("vcm", "bdpm"):
            film, st, _ = rs.render_vcm(ne, mode=R.VCM_MODE_VCM if mode == "vcm" else R.VCM_MODE_BDPM,
                                        num_iteration_pass=passes,
                                        num_photon_trace_samples=int(p.get("num_photon_trace_samples", 10000)),
                                        initial_radius=float(p.get("initial_radius", 0.1)),
                                        alpha=float(p.get("alpha", 0.7)), **common)
        else:
            raise SceneFileError(f"renderer::vcm mode {mode!r} is unknown")
    elif t == "pm":
        # renderer_pm.cpp:61-70; the photon map type (kdtree, naive) finds the same photons
        film, st, _ = rs.render_pm(n, num_photon_trace_samples=int(p.get("num_photon_trace_samples", 100000)),
                                   finalgather=int(p.get("finalgather", 1)), radius=float(p.get("radius", 0.01)),
                                   max_num_vertices=int(p.get("max_num_vertices", -1)), seed=seed, raster_mode=rm)
    elif t == "ppm":
        # renderer_ppm.cpp:76-89: max_num_vertices is required there
        if "max_num_vertices" not in p:
            raise SceneFileError("renderer::ppm needs max_num_vertices")
        film, st, _ = rs.render_ppm(int(p.get("num_samples", 100000)), int(p["max_num_vertices"]),
                                    num_iteration_pass=int(p.get("num_iteration_pass", 1000)),
                                    num_photon_trace_samples=int(p.get("num_photon_trace_samples", 100)),
                                    initial_radius=float(p.get("initial_radius", 0.1)),
                                    alpha=float(p.get("alpha", 0.7)), seed=seed, raster_mode=rm)
    elif t == "sppm":
        # renderer_sppm.cpp:81-96: max_num_vertices is required there; a point per pixel
        if "max_num_vertices" not in p:
            raise SceneFileError("renderer::sppm needs max_num_vertices")
        film, st, _ = rs.render_sppm(int(p["max_num_vertices"]),
                                     num_iteration_pass=int(p.get("num_iteration_pass", 1000)),
                                     num_photon_trace_samples=int(p.get("num_photon_trace_samples", 100)),
                                     initial_radius=float(p.get("initial_radius", 0.1)),
                                     alpha=float(p.get("alpha", 0.7)), seed=seed)
    elif t == "raycast":
        film, st = rs.render_raycast()
    elif t == "ltdirect":
        film, st, _ = rs.render_ltdirect(n, max_num_vertices=int(p.get("max_num_vertices", -1)), seed=seed)
    else:
        raise SceneFileError(f"renderer::{t} is not supported here")
    return film, st


def render_file(path, output, seed=42):
    """main.cpp's `render -s path -o output`: load, render, save the film
    with its type's suffix; returns the written file name."""
    from . import image
    loaded = load_scene(path)
    film, _ = render(loaded, seed=seed)
    ty = {"radiancehdr": image.RADIANCEHDR, "openexr": image.OPENEXR, "png": image.PNG}[loaded["film"]["type"]]
    return image.save(film, output, ty)
