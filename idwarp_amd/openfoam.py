"""OpenFOAM polyMesh (ASCII) input for USMesh.

The reference reads an OpenFOAM case through pyOFM (UnstructuredMesh.py:
1004-1049: volume points, faces, boundary patches) and builds the warping
surface from every boundary patch of type patch, wall or slip
(_computeOFConn, :967-1002): each face contributes its own copy of its
points, in boundary order, and pointReduce merges the copies.  This module
reads the same files directly: constant/polyMesh/{points, faces, boundary}
(``points_orig`` when ``points`` is absent), ASCII format.
"""
import os
import re

import numpy as np

_COMMENT = re.compile(r"//[^\n]*|/\*.*?\*/", re.S)


def _body(path):
    """File text after the FoamFile header, comments removed."""
    with open(path) as f:
        text = _COMMENT.sub(" ", f.read())
    m = re.search(r"FoamFile\s*\{[^}]*\}", text)
    if m:
        header = m.group(0)
        if re.search(r"format\s+binary", header):
            raise ValueError("%s: binary OpenFOAM files are not supported (ASCII only)" % path)
        text = text[m.end():]
    return text


def _list_start(text):
    """(count, index just after the opening parenthesis) of the top-level list."""
    m = re.search(r"(\d+)\s*\(", text)
    if not m:
        raise ValueError("no OpenFOAM list found")
    return int(m.group(1)), m.end()


def read_points(path):
    text = _body(path)
    n, i = _list_start(text)
    vals = np.array(re.findall(r"[-+0-9.eE]+", text[i:text.rindex(")")]), dtype=np.float64)
    if vals.size != 3 * n:
        raise ValueError("%s: expected %d points, found %d values" % (path, n, vals.size))
    return vals.reshape(n, 3)


def read_faces(path):
    """Faces as (faceSizes, flat connectivity) in file order."""
    text = _body(path)
    n, i = _list_start(text)
    toks = re.findall(r"(\d+)\s*\(([^)]*)\)", text[i:])
    if len(toks) != n:
        raise ValueError("%s: expected %d faces, found %d" % (path, n, len(toks)))
    sizes = np.fromiter((int(k) for k, _ in toks), dtype=np.int64, count=n)
    conn = np.array(" ".join(v for _, v in toks).split(), dtype=np.int64)
    if conn.size != sizes.sum():
        raise ValueError("%s: face sizes do not match the point lists" % path)
    return sizes, conn


def read_boundary(path):
    """Ordered list of (name, {type, nFaces, startFace})."""
    text = _body(path)
    _, i = _list_start(text)
    out = []
    for m in re.finditer(r"([A-Za-z_][\w.:-]*)\s*\{([^}]*)\}", text[i:]):
        entries = dict(re.findall(r"(\w+)\s+([^;]+);", m.group(2)))
        out.append((m.group(1), {"type": entries.get("type", "").strip(),
                                 "nFaces": int(entries.get("nFaces", 0)),
                                 "startFace": int(entries.get("startFace", 0))}))
    return out


def read_polymesh(case_dir):
    """Volume points, faces and boundary patches of an OpenFOAM case (the
    case directory or its constant/polyMesh directory)."""
    d = _polymesh_dir(case_dir)
    pts_file = os.path.join(d, "points")
    if not os.path.exists(pts_file):
        pts_file = os.path.join(d, "points_orig")
    x0 = read_points(pts_file)
    sizes, conn = read_faces(os.path.join(d, "faces"))
    return {"x0": x0, "faceSizes": sizes, "faceConn": conn, "faceStart": np.concatenate([[0], np.cumsum(sizes)]),
            "boundaries": read_boundary(os.path.join(d, "boundary"))}


def warp_surface(of, types=("patch", "wall", "slip")):
    """(pts, conn, faceSizes) of the warping surface: the faces of every
    boundary patch whose type is in ``types``, each with its own copy of its
    points (_computeOFConn, UnstructuredMesh.py:967-1002)."""
    x0, start, conn_all = of["x0"], of["faceStart"], of["faceConn"]
    face_ids = []
    for _, b in of["boundaries"]:
        if b["type"].lower() in types and b["nFaces"] > 0:
            face_ids.extend(range(b["startFace"], b["startFace"] + b["nFaces"]))
    face_ids = np.asarray(face_ids, dtype=np.int64)
    sizes = (start[face_ids + 1] - start[face_ids]).astype(np.int32)
    idx = np.concatenate([conn_all[start[f]:start[f + 1]] for f in face_ids]) if len(face_ids) else np.zeros(0, int)
    pts = x0[idx]
    conn = np.arange(len(idx), dtype=np.int32)
    return pts, conn, sizes


_HEADER = """FoamFile
{
    version     2.0;
    format      ascii;
    class       %s;
    location    "constant/polyMesh";
    object      %s;
}
"""


def _polymesh_dir(case_dir):
    d = case_dir
    if os.path.isdir(os.path.join(d, "constant", "polyMesh")):
        d = os.path.join(d, "constant", "polyMesh")
    return d


def write_points(case_dir, x):
    """Overwrite the points file of a case (the case directory or its
    constant/polyMesh) with x (N, 3), as pyOFM's writeVolumeMeshPoints does
    for writeGrid (UnstructuredMesh.py:507-511)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    with open(os.path.join(_polymesh_dir(case_dir), "points"), "w") as f:
        f.write(_HEADER % ("vectorField", "points"))
        f.write("\n%d\n(\n" % len(x))
        f.write("".join("(%.17g %.17g %.17g)\n" % tuple(p) for p in x))
        f.write(")\n")


def write_polymesh(case_dir, x0, face_sizes, face_conn, boundaries):
    """Write points, faces and boundary of an ASCII polyMesh under
    case_dir/constant/polyMesh (the files read_polymesh reads)."""
    d = os.path.join(case_dir, "constant", "polyMesh")
    os.makedirs(d, exist_ok=True)
    x0 = np.asarray(x0, dtype=np.float64).reshape(-1, 3)
    with open(os.path.join(d, "points"), "w") as f:
        f.write(_HEADER % ("vectorField", "points"))
        f.write("\n%d\n(\n" % len(x0))
        f.write("".join("(%.17g %.17g %.17g)\n" % tuple(p) for p in x0))
        f.write(")\n")
    sizes = np.asarray(face_sizes, dtype=np.int64)
    conn = np.asarray(face_conn, dtype=np.int64)
    start = np.concatenate([[0], np.cumsum(sizes)])
    with open(os.path.join(d, "faces"), "w") as f:
        f.write(_HEADER % ("faceList", "faces"))
        f.write("\n%d\n(\n" % len(sizes))
        f.write("".join("%d(%s)\n" % (n, " ".join(map(str, conn[start[i]:start[i] + n]))) for i, n in enumerate(sizes)))
        f.write(")\n")
    with open(os.path.join(d, "boundary"), "w") as f:
        f.write(_HEADER % ("polyBoundaryMesh", "boundary"))
        f.write("\n%d\n(\n" % len(boundaries))
        for name, b in boundaries:
            f.write("    %s\n    {\n        type            %s;\n        nFaces          %d;\n"
                    "        startFace       %d;\n    }\n" % (name, b["type"], b["nFaces"], b["startFace"]))
        f.write(")\n")
