"""OpenFOAM ``constant/polyMesh`` (ASCII) reader and writer for the path's
``Mesh``: the format SEMO loads real cases from
(``SEMO_Mesh_Load::OpenFoam``, src/load/loadOpenFoam.cpp:7-383).

Reading follows the reference loader's rules:

* ``points``: the lines after the first line that starts with ``(``, one
  ``(x y z)`` per line, up to a line that starts with ``)``
  (loadOpenFoam.cpp:40-77);
* ``faces``: ``n(p0 p1 ...)`` per face, possibly spread over several lines;
  the leading count is read and discarded (loadOpenFoam.cpp:79-141);
* ``owner`` / ``neighbour``: one integer per face in order; a negative
  neighbour ends the list (loadOpenFoam.cpp:143-211);
* ``boundary``: one block per patch, its name on the line before ``{``,
  ``nFaces`` and ``startFace`` read from it (loadOpenFoam.cpp:213-336);
* each cell's point list is the first-appearance order of the points of its
  faces, visiting faces in order, internal faces for both cells and boundary
  faces for their owner (``connectCelltoPoints``, src/mesh/build.cpp:642-697).
  That order defines the vertex stencil's order (geometric.cpp:1800-1841).

The writer emits the same subset of the format (``FoamFile`` headers, counts
and lists), enough for this reader and for OpenFOAM's own ASCII reader.
"""
from __future__ import annotations

import os

import numpy as np

from .mesh import Mesh


def _list_lines(path):
    """The lines of a file's main list: after the first line starting with
    '(' up to the first line starting with ')'."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    for k, line in enumerate(lines):
        if line[:1] == "(":
            out = []
            for item in lines[k + 1:]:
                if item[:1] == ")":
                    return out
                out.append(item)
            return out
    raise ValueError(f"{path}: no list found")


def _read_points(path):
    pts = []
    for line in _list_lines(path):
        s = line.replace("(", " ").replace(")", " ").split()
        if s:
            pts.append([float(v) for v in s[:3]])
    return np.asarray(pts, dtype=np.float64).reshape(-1, 3)


def _read_faces(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    start = next(k for k, line in enumerate(lines) if line[:1] == "(")
    faces, pending = [], ""
    for line in lines[start + 1:]:
        if line[:1] == ")" and not pending:
            break
        if len(line) == 1 and line != ")":
            continue
        pending += " " + line.rstrip()
        if ")" not in line:
            continue
        s = pending.replace("(", " ", 1).replace(")", " ", 1).split()
        faces.append([int(v) for v in s[1:]])       # s[0] is the count
        pending = ""
    return faces


def _read_ints(path, stop_negative=False):
    out = []
    for line in _list_lines(path):
        for v in line.split():
            i = int(v)
            if stop_negative and i < 0:
                return out
            out.append(i)
    return out


def _read_boundary(path):
    items = _list_lines(path)
    patches, name, block, back, inside = [], None, [], "", False
    for item in items:
        if inside:
            if "}" in item:
                nf = sf = None
                for b in block:
                    t = b.replace(";", " ").split()
                    if "nFaces" in b and len(t) > 1:
                        nf = int(t[1])
                    if "startFace" in b and len(t) > 1:
                        sf = int(t[1])
                patches.append((name.strip(), sf, nf))
                inside, block = False, []
            else:
                block.append(item)
        elif "{" in item:
            name, inside, block = back, True, []
        back = item
    return patches


def read_polymesh(folder: str) -> Mesh:
    """Mesh from an OpenFOAM ASCII ``polyMesh`` folder (points, faces,
    owner, neighbour, boundary)."""
    pts = _read_points(os.path.join(folder, "points"))
    faces = _read_faces(os.path.join(folder, "faces"))
    owner = np.asarray(_read_ints(os.path.join(folder, "owner")), dtype=np.int32)
    nbr = np.asarray(_read_ints(os.path.join(folder, "neighbour"), stop_negative=True), dtype=np.int32)
    patches = _read_boundary(os.path.join(folder, "boundary"))
    nf, nif = len(faces), len(nbr)
    if len(owner) != nf:
        raise ValueError(f"{folder}: {len(owner)} owners for {nf} faces")
    nc = int(max(owner.max(initial=-1), nbr.max(initial=-1))) + 1
    face_ptr = np.zeros(nf + 1, dtype=np.int32)
    face_ptr[1:] = np.cumsum([len(f) for f in faces])
    face_pts = np.asarray([p for f in faces for p in f], dtype=np.int32)
    cells = [[] for _ in range(nc)]
    seen = [set() for _ in range(nc)]

    def add(c, f):
        for p in f:
            if p not in seen[c]:
                seen[c].add(p)
                cells[c].append(p)
    for k, f in enumerate(faces):
        add(int(owner[k]), f)
        if k < nif:
            add(int(nbr[k]), f)
    cell_ptr = np.zeros(nc + 1, dtype=np.int32)
    cell_ptr[1:] = np.cumsum([len(c) for c in cells])
    cell_pts = np.asarray([p for c in cells for p in c], dtype=np.int32)
    return Mesh(points=pts, owner=owner, neighbour=nbr, face_ptr=face_ptr, face_pts=face_pts,
                cell_ptr=cell_ptr, cell_pts=cell_pts, patch_names=[p[0] for p in patches],
                patch_start=np.asarray([p[1] for p in patches], dtype=np.int32),
                patch_nfaces=np.asarray([p[2] for p in patches], dtype=np.int32))


def _header(cls, obj, note=""):
    return ("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
            f"    class       {cls};\n" + (f"    note        \"{note}\";\n" if note else "") +
            f"    location    \"constant/polyMesh\";\n    object      {obj};\n}}\n\n")


def write_polymesh(mesh: Mesh, folder: str) -> None:
    """The mesh as an OpenFOAM ASCII ``polyMesh`` folder; boundary patches of
    type ``patch``."""
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "points"), "w") as fh:
        fh.write(_header("vectorField", "points"))
        fh.write(f"{mesh.n_points}\n(\n")
        for x, y, z in mesh.points.tolist():      # repr of a float round-trips exactly
            fh.write(f"({x!r} {y!r} {z!r})\n")
        fh.write(")\n")
    with open(os.path.join(folder, "faces"), "w") as fh:
        fh.write(_header("faceList", "faces"))
        fh.write(f"{mesh.n_faces}\n(\n")
        for f in range(mesh.n_faces):
            p = mesh.face_pts[mesh.face_ptr[f]:mesh.face_ptr[f + 1]]
            fh.write(f"{len(p)}(" + " ".join(str(int(v)) for v in p) + ")\n")
        fh.write(")\n")
    note = f"nPoints:{mesh.n_points} nCells:{mesh.n_cells} nFaces:{mesh.n_faces} " \
           f"nInternalFaces:{mesh.n_internal_faces}"
    for name, arr in (("owner", mesh.owner), ("neighbour", mesh.neighbour)):
        with open(os.path.join(folder, name), "w") as fh:
            fh.write(_header("labelList", name, note))
            fh.write(f"{len(arr)}\n(\n")
            fh.write("\n".join(str(int(v)) for v in arr))
            fh.write("\n)\n")
    with open(os.path.join(folder, "boundary"), "w") as fh:
        fh.write(_header("polyBoundaryMesh", "boundary"))
        fh.write(f"{len(mesh.patch_names)}\n(\n")
        for b, name in enumerate(mesh.patch_names):
            fh.write(f"    {name}\n    {{\n        type            patch;\n"
                     f"        nFaces          {int(mesh.patch_nfaces[b])};\n"
                     f"        startFace       {int(mesh.patch_start[b])};\n    }}\n")
        fh.write(")\n")
