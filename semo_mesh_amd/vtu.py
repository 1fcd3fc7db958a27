"""SEMO's ``plot.<rank>.vtu`` (ASCII variant) for the path's ``Mesh`` and
cell fields: the file SEMO saves and restarts from.

Layout (``SEMO_Save::vtu``, src/save/save.cpp:48-767): a VTK unstructured
grid with ``TimeValue`` field data, the cell fields as ``CellData``, the
points (``NodeCoordinates``), the cells as VTK polyhedra (``connectivity``,
``offsets``, ``types`` = 42, the per-cell face stream ``faces`` with each
cell's faces in ascending face order, ``faceoffsets``), then, after the
piece, ``owner`` and ``neighbour`` per face (-1 on boundary faces) and the
patches as ``bcName`` / ``bcStartFace`` / ``bcNFaces`` / ``bcNeighbProcNo``.

Reading follows ``SEMO_Mesh_Load::vtu`` (src/load/load.cpp:114-837): cell
points from connectivity/offsets, each cell's faces in ascending face order,
and each face's points from the first cell whose stream lists it
(load.cpp:352-413).  Only the ASCII format is handled; SEMO's binary and
zlib/base64 variants carry the same arrays.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

from .mesh import Mesh


def _cell_faces(mesh: Mesh):
    """Each cell's faces in ascending face order (load.cpp:373-385)."""
    cf = [[] for _ in range(mesh.n_cells)]
    nif = mesh.n_internal_faces
    for f in range(mesh.n_faces):
        cf[int(mesh.owner[f])].append(f)
        if f < nif:
            cf[int(mesh.neighbour[f])].append(f)
    return cf


def _array(out, typ, name, values, ncomp=1, fmt=repr):
    comp = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
    out.append(f'     <DataArray type="{typ}" Name="{name}"{comp} format="ascii">')
    out.append(" ".join(fmt(v) for v in values))
    out.append("     </DataArray>")


def write_vtu(path: str, mesh: Mesh, fields: dict | None = None, time: float = 0.0) -> None:
    """Mesh and cell fields (name -> [n] or [n, 3]) as SEMO's ASCII VTU."""
    fields = fields or {}
    nc = mesh.n_cells
    out = ['<?xml version="1.0"?>',
           ' <VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
           '  <UnstructuredGrid>', '    <FieldData>']
    _array(out, "Float64", "TimeValue", [float(time)])
    out += ['    </FieldData>', f'   <Piece NumberOfPoints="{mesh.n_points}" NumberOfCells="{nc}">',
            '    <PointData>', '    </PointData>', '    <CellData>']
    for name, v in fields.items():
        a = np.asarray(v, dtype=np.float64)
        _array(out, "Float64", name, a.reshape(-1).tolist(), 3 if a.ndim == 2 else 1)
    out += ['    </CellData>', '    <Points>']
    _array(out, "Float64", "NodeCoordinates", mesh.points.reshape(-1).tolist(), 3)
    out += ['    </Points>', '   <Cells>']
    _array(out, "Int32", "connectivity", mesh.cell_pts.tolist(), fmt=str)
    _array(out, "Int32", "offsets", mesh.cell_ptr[1:].tolist(), fmt=str)
    _array(out, "Int32", "types", [42] * nc, fmt=str)
    stream, offs = [], []
    for faces in _cell_faces(mesh):
        stream.append(len(faces))
        for f in faces:
            p = mesh.face_pts[mesh.face_ptr[f]:mesh.face_ptr[f + 1]].tolist()
            stream.append(len(p))
            stream += p
        offs.append(len(stream))
    _array(out, "Int32", "faces", stream, fmt=str)
    _array(out, "Int32", "faceoffsets", offs, fmt=str)
    out += ['   </Cells>', '  </Piece>']
    nbr = np.full(mesh.n_faces, -1, dtype=np.int64)
    nbr[:mesh.n_internal_faces] = mesh.neighbour
    _array(out, "Int32", "owner", mesh.owner.tolist(), fmt=str)
    _array(out, "Int32", "neighbour", nbr.tolist(), fmt=str)
    out.append('     <DataArray type="Char" Name="bcName" format="ascii">')
    out.append(" ".join(n.strip() for n in mesh.patch_names))
    out.append("     </DataArray>")
    _array(out, "Int32", "bcStartFace", [int(v) for v in mesh.patch_start], fmt=str)
    _array(out, "Int32", "bcNFaces", [int(v) for v in mesh.patch_nfaces], fmt=str)
    _array(out, "Int32", "bcNeighbProcNo", [-1] * len(mesh.patch_names), fmt=str)
    out += ['  </UnstructuredGrid>', ' </VTKFile>']
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def read_vtu(path: str):
    """(Mesh, {field name: [n] or [n, 3]}, time) from SEMO's ASCII VTU."""
    root = ET.parse(path).getroot()
    arrays, ncomp, kinds = {}, {}, {}
    for parent in root.iter():
        for da in parent.findall("DataArray"):
            name = da.get("Name")
            text = (da.text or "").split()
            if da.get("format", "ascii") != "ascii":
                raise ValueError(f"{path}: only the ascii format is read ({name})")
            arrays[name] = text
            ncomp[name] = int(da.get("NumberOfComponents", "1"))
            kinds[name] = parent.tag
    ints = lambda n: np.asarray([int(v) for v in arrays[n]], dtype=np.int64)
    pts = np.asarray([float(v) for v in arrays["NodeCoordinates"]]).reshape(-1, 3)
    conn, offs = ints("connectivity"), ints("offsets")
    owner, nbr_all = ints("owner"), ints("neighbour")
    nc = len(offs)
    nif = int(np.count_nonzero(nbr_all != -1))
    if not np.all(nbr_all[:nif] != -1):
        raise ValueError(f"{path}: internal faces must come first")
    # cells' faces in ascending face order, then the face points from the
    # first cell whose stream lists the face (load.cpp:373-413)
    cf = [[] for _ in range(nc)]
    for f in range(len(owner)):
        cf[int(owner[f])].append(f)
        if nbr_all[f] != -1:
            cf[int(nbr_all[f])].append(f)
    stream, foffs = ints("faces"), ints("faceoffsets")
    fpts = [None] * len(owner)
    start = 0
    for c in range(nc):
        k = start
        nfc = int(stream[k]); k += 1
        for j in range(nfc):
            npt = int(stream[k]); k += 1
            f = cf[c][j]
            if fpts[f] is None:
                fpts[f] = stream[k:k + npt].tolist()
            k += npt
        start = int(foffs[c])
    face_ptr = np.zeros(len(owner) + 1, dtype=np.int32)
    face_ptr[1:] = np.cumsum([len(p) for p in fpts])
    cell_ptr = np.zeros(nc + 1, dtype=np.int32)
    cell_ptr[1:] = offs
    mesh = Mesh(points=pts, owner=owner.astype(np.int32), neighbour=nbr_all[:nif].astype(np.int32),
                face_ptr=face_ptr, face_pts=np.asarray([p for fp in fpts for p in fp], dtype=np.int32),
                cell_ptr=cell_ptr, cell_pts=conn.astype(np.int32),
                patch_names=list(arrays.get("bcName", [])),
                patch_start=ints("bcStartFace").astype(np.int32), patch_nfaces=ints("bcNFaces").astype(np.int32))
    fields = {}
    for name, text in arrays.items():
        if kinds[name] == "CellData":
            a = np.asarray([float(v) for v in text])
            fields[name] = a.reshape(-1, 3) if ncomp[name] == 3 else a
    time = float(arrays["TimeValue"][0]) if "TimeValue" in arrays else 0.0
    return mesh, fields, time
