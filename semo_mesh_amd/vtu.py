"""SEMO's ``plot.<rank>.vtu`` for the path's ``Mesh`` and cell fields: the
file SEMO saves and restarts from, in its three encodings.

Layout (``SEMO_Mesh_Save::vtu``, src/save/save.cpp:48-767): a VTK
unstructured grid with ``TimeValue`` field data, ``pointLevels`` point data,
the cell fields as ``CellData`` (``pressure``, ``velocity`` (3 components),
``temperature``, ``volume-fraction-<species 0>``, ``mass-fraction-<species
0>``, ``cellLevels``, ``cellGroups``), the points (``NodeCoordinates``), the
cells as VTK polyhedra (``connectivity``, ``offsets``, ``types`` = 42, the
per-cell face stream ``faces`` with each cell's faces in ascending face
order, ``faceoffsets``), then, after the grid, ``owner`` and ``neighbour``
per face (-1 on boundary faces) and the patches as ``bcName`` /
``bcStartFace`` / ``bcNFaces`` / ``bcNeighbProcNo`` (always ascii).

Encodings (``saveFormat`` / ``saveCompression`` of controlDict):

* ``ascii``: the values separated by spaces (save/saveAscii.cpp);
* ``binary``: one base64 string of an 8-byte little-endian header and the
  raw little-endian values (save/saveBinary.cpp).  SEMO's header is 8 times
  the number of values whatever the value size, and its reader takes the
  count from it (load/loadBinary.cpp:38-43); standard VTK writes the byte
  count.  Reading accepts either; writing follows SEMO, so that SEMO reads
  the file back;
* ``binary`` with ``compressor="vtkZLibDataCompressor"`` on the VTKFile
  element: a base64 header of UInt64 [blocks, block size, last block size,
  compressed size of each block] followed by the base64 of the zlib streams
  (save/saveCompress.cpp: one block; load/loadCompress.cpp reads a 44-char
  one-block header).  Reading accepts several blocks; writing uses one.

Reading follows ``SEMO_Mesh_Load::vtu`` (src/load/load.cpp:114-837): cell
points from connectivity/offsets, each cell's faces in ascending face order,
and each face's points from the first cell whose stream lists it
(load.cpp:352-413).
"""
from __future__ import annotations

import base64
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

from .mesh import Mesh

_DTYPES = {"Float64": "<f8", "Float32": "<f4", "Int32": "<i4", "Int64": "<i8", "UInt8": "<u1",
           "Int8": "<i1", "UInt32": "<u4", "UInt64": "<u8"}


def _cell_faces(mesh: Mesh):
    """Each cell's faces in ascending face order (load.cpp:373-385)."""
    cf = [[] for _ in range(mesh.n_cells)]
    nif = mesh.n_internal_faces
    for f in range(mesh.n_faces):
        cf[int(mesh.owner[f])].append(f)
        if f < nif:
            cf[int(mesh.neighbour[f])].append(f)
    return cf


# ------------------------------------------------------------------ encoding --
def _encode_binary(a: np.ndarray) -> str:
    """save/saveBinary.cpp: base64 of [UInt64 8 * count][values]."""
    hdr = struct.pack("<Q", 8 * a.size)
    return base64.b64encode(hdr + a.tobytes()).decode("ascii")


def _encode_compressed(a: np.ndarray, level: int) -> str:
    """save/saveCompress.cpp: base64 of [1, raw bytes, raw bytes, zlib
    bytes] (UInt64), then base64 of the zlib stream of the raw values."""
    raw = a.tobytes()
    z = zlib.compress(raw, level)
    hdr = struct.pack("<4Q", 1, len(raw), len(raw), len(z))
    return base64.b64encode(hdr).decode("ascii") + base64.b64encode(z).decode("ascii")


def _decode_binary(text: str, dtype: str, name: str) -> np.ndarray:
    raw = base64.b64decode(text)
    size = np.dtype(dtype).itemsize
    if len(raw) < 8 or (len(raw) - 8) % size:
        raise ValueError(f"binary array {name}: {len(raw)} bytes do not hold a header and whole values")
    n = (len(raw) - 8) // size
    (h,) = struct.unpack("<Q", raw[:8])
    if h not in (8 * n, size * n):
        raise ValueError(f"binary array {name}: header {h} does not match {n} values")
    return np.frombuffer(raw, dtype=dtype, count=n, offset=8).copy()


def _decode_compressed(text: str, dtype: str, name: str) -> np.ndarray:
    nb = struct.unpack("<Q", base64.b64decode(text[:12])[:8])[0]
    hbytes = 8 * (3 + nb)
    hchars = 4 * ((hbytes + 2) // 3)
    hdr = struct.unpack(f"<{3 + nb}Q", base64.b64decode(text[:hchars])[:hbytes])
    bsize, last = hdr[1], hdr[2]
    data = base64.b64decode(text[hchars:])
    out, pos = [], 0
    for b in range(nb):
        c = hdr[3 + b]
        blk = zlib.decompress(data[pos:pos + c])
        want = last if (b == nb - 1 and last) else bsize
        if len(blk) != want:
            raise ValueError(f"compressed array {name}: block {b} inflates to {len(blk)} bytes, not {want}")
        out.append(blk)
        pos += c
    raw = b"".join(out)
    if len(raw) % np.dtype(dtype).itemsize:
        raise ValueError(f"compressed array {name}: {len(raw)} bytes are not whole values")
    return np.frombuffer(raw, dtype=dtype).copy()


class _Writer:
    def __init__(self, fmt: str, compression: int):
        if fmt not in ("ascii", "binary"):
            raise ValueError("fmt must be 'ascii' or 'binary'")
        self.fmt, self.level = fmt, int(compression)
        self.lines: list[str] = []

    def array(self, indent: str, typ: str, name: str, values, ncomp: int = 1, extra: str = "",
              force_ascii: bool = False) -> None:
        a = np.ascontiguousarray(values, dtype=_DTYPES[typ]).reshape(-1)
        fmt = "ascii" if force_ascii else self.fmt
        comp = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
        self.lines.append(f'{indent}<DataArray type="{typ}"{extra} Name="{name}"{comp} format="{fmt}">')
        if fmt == "ascii":
            conv = repr if a.dtype.kind == "f" else str
            self.lines.append(" ".join(conv(v.item()) for v in a) + " ")
        elif self.level:
            self.lines.append(_encode_compressed(a, self.level))
        else:
            self.lines.append(_encode_binary(a))
        self.lines.append(f"{indent}</DataArray>")


def semo_fields(P, U, V, W, VF=None, T=None, species: str = "water") -> dict:
    """The cell arrays SEMO writes and reads back, under its names:
    pressure, velocity (n x 3), temperature (300 K when not given),
    volume-fraction-<species> and mass-fraction-<species> (both VF, as
    initial.cpp sets them; 1 when not given)."""
    n = len(P)
    VF = np.ones(n) if VF is None else np.asarray(VF, dtype=np.float64)
    return {"pressure": np.asarray(P, dtype=np.float64),
            "velocity": np.stack([np.asarray(U), np.asarray(V), np.asarray(W)], axis=1).astype(np.float64),
            "temperature": np.full(n, 300.0) if T is None else np.asarray(T, dtype=np.float64),
            f"volume-fraction-{species}": VF, f"mass-fraction-{species}": VF.copy()}


def write_vtu(path: str, mesh: Mesh, fields: dict | None = None, time: float = 0.0, fmt: str = "ascii",
              compression: int = 0, point_levels=None, cell_levels=None, cell_groups=None) -> None:
    """Mesh and cell fields (name -> [n] or [n, 3]) as SEMO's VTU, in
    SEMO's array order.  fmt 'ascii' (shortest round-trip decimals, exact)
    or 'binary'; compression > 0 (zlib level) with 'binary' gives the
    compressed encoding.  Levels and groups default to zero.  SEMO's loader
    needs pressure, velocity and temperature among the fields (semo_fields
    gives them)."""
    fields = fields or {}
    nc, npt = mesh.n_cells, mesh.n_points
    if compression and fmt != "binary":
        raise ValueError("compression needs fmt='binary'")
    w = _Writer(fmt, compression)
    comp_attr = ' compressor="vtkZLibDataCompressor"' if (fmt == "binary" and compression) else ""
    L = w.lines
    L += ['<?xml version="1.0"?>',
          f' <VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64"'
          f'{comp_attr}>', '  <UnstructuredGrid>', '    <FieldData>']
    w.array("     ", "Float64", "TimeValue", [float(time)], extra=' NumberOfTuples="1"')
    L += ['    </FieldData>', f'   <Piece NumberOfPoints="{npt}" NumberOfCells="{nc}">', '    <PointData>']
    w.array("     ", "Int32", "pointLevels", np.zeros(npt, np.int32) if point_levels is None else point_levels)
    L += ['    </PointData>', '    <CellData>']
    for name, v in fields.items():
        a = np.asarray(v, dtype=np.float64)
        w.array("     ", "Float64", name, a, 3 if a.ndim == 2 else 1)
    w.array("     ", "Int32", "cellLevels", np.zeros(nc, np.int32) if cell_levels is None else cell_levels)
    w.array("     ", "Int32", "cellGroups", np.zeros(nc, np.int32) if cell_groups is None else cell_groups)
    L += ['    </CellData>', '    <Points>']
    w.array("     ", "Float64", "NodeCoordinates", mesh.points, 3)
    L += ['    </Points>', '   <Cells>']
    w.array("    ", "Int32", "connectivity", mesh.cell_pts)
    w.array("    ", "Int32", "offsets", mesh.cell_ptr[1:])
    w.array("    ", "Int32", "types", np.full(nc, 42, np.int32))
    stream, offs = [], []
    for faces in _cell_faces(mesh):
        stream.append(len(faces))
        for f in faces:
            p = mesh.face_pts[mesh.face_ptr[f]:mesh.face_ptr[f + 1]].tolist()
            stream.append(len(p))
            stream += p
        offs.append(len(stream))
    w.array("    ", "Int32", "faces", stream, extra=' IdType="1"')
    w.array("    ", "Int32", "faceoffsets", offs, extra=' IdType="1"')
    L += ['   </Cells>', '  </Piece>', ' </UnstructuredGrid>']
    nbr = np.full(mesh.n_faces, -1, dtype=np.int64)
    nbr[:mesh.n_internal_faces] = mesh.neighbour
    w.array(" ", "Int32", "owner", mesh.owner)
    w.array(" ", "Int32", "neighbour", nbr)
    L.append(' <DataArray type="Char" Name="bcName" format="ascii">')
    L.append(" ".join(n.strip() for n in mesh.patch_names) + " ")
    L.append(" </DataArray>")
    w.array(" ", "Int32", "bcStartFace", [int(v) for v in mesh.patch_start], force_ascii=True)
    w.array(" ", "Int32", "bcNFaces", [int(v) for v in mesh.patch_nfaces], force_ascii=True)
    w.array(" ", "Int32", "bcNeighbProcNo", [-1] * len(mesh.patch_names), force_ascii=True)
    L.append('</VTKFile>')
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def write_grid_vtu(path: str, mesh: Mesh) -> None:
    """The mesh alone in SEMO's grid layout (``SEMO_Mesh_Save::vtu(mesh)``,
    src/save/saveOnlyMesh.cpp:8-147): what mesh converters write to
    ./grid/0/ and what the initial-field utility reads (initial.cpp:1498):
    ascii, one point per line (here to 17 significant digits), each cell's
    points and each cell's face stream on lines of their own, and the face
    and patch lists in elements ``owner``, ``neighbour``, ``bcName``,
    ``bcStartFace``, ``bcNFaces``, ``bcNeighbProcNo`` after the grid."""
    nc = mesh.n_cells
    cptr, cpts = np.asarray(mesh.cell_ptr), np.asarray(mesh.cell_pts)
    L = ['<?xml version="1.0"?>',
         ' <VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
         '  <UnstructuredGrid>', f'   <Piece NumberOfPoints="{mesh.n_points}" NumberOfCells="{nc}">',
         '    <PointData>', '    </PointData>', '    <CellData>', '    </CellData>', '    <Points>',
         '     <DataArray type="Float64" Name="NodeCoordinates" NumberOfComponents="3" format="ascii">']
    L += [f"{x!r} {y!r} {z!r}" for x, y, z in np.asarray(mesh.points, dtype=np.float64).tolist()]
    L += ['    </DataArray>', '   </Points>', '   <Cells>',
          '    <DataArray type="Int64" Name="connectivity" format="ascii">']
    L += [" ".join(str(int(p)) for p in cpts[cptr[c]:cptr[c + 1]]) + " " for c in range(nc)]
    L += ['    </DataArray>', '    <DataArray type="Int64" Name="offsets" format="ascii">',
          " ".join(str(int(v)) for v in cptr[1:]) + " ", '    </DataArray>',
          '    <DataArray type="UInt8" Name="types" format="ascii">', "42 " * nc, '    </DataArray>',
          '    <DataArray type="Int64" IdType="1" Name="faces" format="ascii">']
    offs, off = [], 0
    for faces in _cell_faces(mesh):
        L.append(str(len(faces)))
        num = 1 + len(faces)
        for f in faces:
            p = mesh.face_pts[mesh.face_ptr[f]:mesh.face_ptr[f + 1]].tolist()
            L.append(f"{len(p)} " + " ".join(str(int(q)) for q in p) + " ")
            num += len(p)
        off += num
        offs.append(off)
    L += ['    </DataArray>', '    <DataArray type="Int64" IdType="1" Name="faceoffsets" format="ascii">',
          " ".join(str(v) for v in offs) + " ", '    </DataArray>', '   </Cells>', '  </Piece>',
          ' </UnstructuredGrid>']
    nbr = np.full(mesh.n_faces, -1, dtype=np.int64)
    nbr[:mesh.n_internal_faces] = mesh.neighbour
    for tag, vals in (("owner", mesh.owner), ("neighbour", nbr), ("bcName", [n.strip() for n in mesh.patch_names]),
                      ("bcStartFace", mesh.patch_start), ("bcNFaces", mesh.patch_nfaces),
                      ("bcNeighbProcNo", [-1] * len(mesh.patch_names))):
        L += [f" <{tag}>", " ".join(str(v if isinstance(v, str) else int(v)) for v in vals) + " ", f" </{tag}>"]
    L.append('</VTKFile>')
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def read_vtu_arrays(path: str):
    """{name: (values, components, parent element tag)} of every DataArray
    of SEMO's VTU, decoded from its encoding (ascii, binary or compressed)."""
    root = ET.parse(path).getroot()
    compressed = root.get("compressor") == "vtkZLibDataCompressor"
    if root.get("compressor") not in (None, "vtkZLibDataCompressor"):
        raise ValueError(f"{path}: compressor {root.get('compressor')} is not read")
    if root.get("header_type", "UInt64") != "UInt64":
        raise ValueError(f"{path}: only UInt64 headers are read")
    out = {}
    for parent in root.iter():
        for da in parent.findall("DataArray"):
            name, typ = da.get("Name"), da.get("type")
            fmt = da.get("format", "ascii")
            text = (da.text or "").strip()
            if typ == "Char" or typ == "String":
                vals = text.split()
            elif fmt == "ascii":
                dt = np.dtype(_DTYPES[typ])
                vals = np.asarray([float(v) for v in text.split()] if dt.kind == "f"
                                  else [int(v) for v in text.split()], dtype=dt)
            elif fmt == "binary":
                tok = "".join(text.split())
                vals = (_decode_compressed(tok, _DTYPES[typ], name) if compressed
                        else _decode_binary(tok, _DTYPES[typ], name))
            else:
                raise ValueError(f"{path}: format {fmt} of {name} is not read")
            out[name] = (vals, int(da.get("NumberOfComponents", "1")), parent.tag)
    # SEMO's grid layout keeps the face and patch lists in elements of their own
    for tag in ("owner", "neighbour", "bcName", "bcStartFace", "bcNFaces", "bcNeighbProcNo"):
        el = root.find(tag)
        if tag not in out and el is not None:
            toks = (el.text or "").split()
            out[tag] = (toks if tag == "bcName" else np.asarray([int(t) for t in toks], dtype=np.int64), 1, "VTKFile")
    return out


def read_vtu(path: str, with_levels: bool = False):
    """(Mesh, {field name: [n] or [n, 3]}, time) from SEMO's VTU in any of
    its encodings; with_levels: also {"pointLevels", "cellLevels",
    "cellGroups"} as a fourth item."""
    arrays = read_vtu_arrays(path)
    get = lambda n: np.asarray(arrays[n][0])     # noqa: E731
    ints = lambda n: get(n).astype(np.int64)     # noqa: E731
    pts = get("NodeCoordinates").astype(np.float64).reshape(-1, 3)
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
                patch_names=list(arrays["bcName"][0]) if "bcName" in arrays else [],
                patch_start=ints("bcStartFace").astype(np.int32), patch_nfaces=ints("bcNFaces").astype(np.int32))
    fields = {}
    for name, (vals, ncomp, kind) in arrays.items():
        if kind == "CellData" and name not in ("cellLevels", "cellGroups"):
            a = np.asarray(vals, dtype=np.float64)
            fields[name] = a.reshape(-1, ncomp) if ncomp > 1 else a
    time = float(arrays["TimeValue"][0][0]) if "TimeValue" in arrays else 0.0
    if with_levels:
        lv = {n: ints(n).astype(np.int32) for n in ("pointLevels", "cellLevels", "cellGroups") if n in arrays}
        return mesh, fields, time, lv
    return mesh, fields, time
