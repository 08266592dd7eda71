"""trianglemesh::raw (src/liblightmetrica/asset/trianglemesh/
trianglemesh_raw.cpp:41-82): a mesh given inline as whitespace-separated
number lists under `positions`, `normals`, `texcoords` and `faces`.

Lists are read as PropertyNode::As<std::vector<Float>> /
<std::vector<unsigned int>> read them (property.h:203-220): numbers in order
until the first token that is not one, floats parsed in double precision and
stored as float32.  Validation follows TriangleMesh_Raw::Load: `positions`
is required and a multiple of 3; `normals`, when present, as many as the
positions; `texcoords`, when present, two per vertex; `faces` (absent: no
faces) a multiple of 3.
"""
import re

import numpy as np

_UINT = re.compile(r"^\+?[0-9]+$")


def _floats(s):
    out = []
    for tok in str(s).split():
        try:
            out.append(float(tok))
        except ValueError:
            break
    return np.array(out, np.float64).astype(np.float32)


def _uints(s):
    out = []
    for tok in str(s).split():
        if not _UINT.match(tok) or int(tok) > 0xFFFFFFFF:
            break
        out.append(int(tok))
    return np.array(out, np.uint32)


def load_raw_mesh(node):
    """`node`: the asset's params as a mapping, or YAML text of one.  Returns
    dict(positions (V, 3), normals (V, 3) or None, texcoords (V, 2) or None,
    faces (F, 3) uint32)."""
    if isinstance(node, str):
        import yaml
        node = yaml.safe_load(node)
    if not isinstance(node, dict) or node.get("positions") is None:
        raise ValueError("trianglemesh::raw: missing 'positions'")
    ps = _floats(node["positions"])
    if ps.size % 3:
        raise ValueError(f"Invalid number of elements in 'positions': {ps.size}")
    ns = ts = None
    if node.get("normals") is not None:
        ns = _floats(node["normals"])
        if ns.size != ps.size:
            raise ValueError(f"Invalid number of elements in 'normals': {ns.size}")
    if node.get("texcoords") is not None:
        ts = _floats(node["texcoords"])
        if ts.size != ps.size // 3 * 2:
            raise ValueError(f"Invalid number of elements in 'texcoords': {ts.size}")
    fs = _uints(node["faces"]) if node.get("faces") is not None else np.zeros(0, np.uint32)
    if fs.size % 3:
        raise ValueError(f"Invalid number of elements in 'faces': {fs.size}")
    return {"positions": ps.reshape(-1, 3), "normals": None if ns is None else ns.reshape(-1, 3),
            "texcoords": None if ts is None else ts.reshape(-1, 2), "faces": fs.reshape(-1, 3)}


def dump_raw_mesh(positions, faces, normals=None, texcoords=None):
    """YAML text of a trianglemesh::raw asset's params (folded number lists,
    one vertex or face per line, floats with 9 significant digits so that
    float32 values read back exactly)."""
    def block(a, fmt):
        return ">\n" + "".join("  " + " ".join(fmt % v for v in row) + "\n" for row in np.asarray(a))
    out = "positions: " + block(positions, "%.9g")
    if normals is not None:
        out += "normals: " + block(normals, "%.9g")
    if texcoords is not None:
        out += "texcoords: " + block(texcoords, "%.9g")
    out += "faces: " + block(faces, "%d")
    return out
