"""Synthetic unstructured meshes in SEMO's face-based layout.

The arrays produced here are what SEMO's `SEMO_Mesh_Load::vtu` builds from a
`plot.<rank>.vtu` file (src/load/load.cpp:352-438):

* ``owner[f]`` / ``neighbour[f]`` per face, internal faces first in
  upper-triangular order (sorted by owner, then by neighbour), boundary faces
  after them grouped per patch;
* ``face_pts`` (CSR by ``face_ptr``): the face's points, ordered so that the
  right-hand normal points from owner to neighbour (outwards on boundaries);
* ``cell_pts`` (CSR by ``cell_ptr``): the cell's points in VTK order; SEMO
  reads them from the VTU connectivity (load.cpp:359-369) and the vertex
  stencil is built from this order (src/mesh/geometric.cpp:1800-1841);
* patches as (name, startFace, nFaces) triples (load.cpp:420-438).

Interior vertices may be jittered by a counter-based hash, so that the same
seed gives bit-identical meshes on every host:
``u = splitmix64(seed ^ (3*vertex_id + axis)) >> 11`` scaled to [0, 1), and
the vertex moves by ``(2u - 1) * amp * h_axis``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

DEFAULT_SEED = 20261202

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """Vectorised splitmix64 finaliser on uint64 (wrap-around arithmetic)."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def unit_uniform(seed: int, key: np.ndarray) -> np.ndarray:
    """U[0,1) doubles from splitmix64(seed ^ key), 53 random bits each."""
    z = splitmix64(np.uint64(seed) ^ np.asarray(key, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Mesh:
    points: np.ndarray            # (n_points, 3) float64
    owner: np.ndarray             # (n_faces,) int32
    neighbour: np.ndarray         # (n_internal_faces,) int32
    face_ptr: np.ndarray          # (n_faces+1,) int32
    face_pts: np.ndarray          # int32
    cell_ptr: np.ndarray          # (n_cells+1,) int32
    cell_pts: np.ndarray          # int32
    patch_names: list = field(default_factory=list)
    patch_start: np.ndarray = None  # (n_patches,) int32
    patch_nfaces: np.ndarray = None  # (n_patches,) int32
    shape: tuple = None

    @property
    def n_cells(self) -> int:
        return len(self.cell_ptr) - 1

    @property
    def n_faces(self) -> int:
        return len(self.owner)

    @property
    def n_internal_faces(self) -> int:
        return len(self.neighbour)

    @property
    def n_points(self) -> int:
        return len(self.points)

    def patch(self, name: str) -> int:
        return self.patch_names.index(name)

    def save(self, path: str) -> None:
        np.savez(path, points=self.points, owner=self.owner,
                 neighbour=self.neighbour, face_ptr=self.face_ptr,
                 face_pts=self.face_pts, cell_ptr=self.cell_ptr,
                 cell_pts=self.cell_pts,
                 patch_names=np.array(self.patch_names),
                 patch_start=self.patch_start,
                 patch_nfaces=self.patch_nfaces)

    def write_raw(self, path: str) -> None:
        """Flat little-endian dump read by the C++ drivers:
        header int32[6] = n_points, n_cells, n_faces, n_internal, n_face_pts,
        n_cell_pts, n_patches; then the arrays in declaration order; then
        per patch: int32 start, int32 nfaces, int32 name length, name bytes."""
        with open(path, "wb") as fh:
            hdr = np.array([self.n_points, self.n_cells, self.n_faces,
                            self.n_internal_faces, len(self.face_pts),
                            len(self.cell_pts), len(self.patch_names)],
                           dtype=np.int32)
            hdr.tofile(fh)
            np.ascontiguousarray(self.points, dtype=np.float64).tofile(fh)
            for a in (self.owner, self.neighbour, self.face_ptr, self.face_pts,
                      self.cell_ptr, self.cell_pts):
                np.ascontiguousarray(a, dtype=np.int32).tofile(fh)
            for i, nm in enumerate(self.patch_names):
                b = nm.encode()
                np.array([self.patch_start[i], self.patch_nfaces[i], len(b)],
                         dtype=np.int32).tofile(fh)
                fh.write(b)


def _quad_faces_plus(pid, axis, i, j, k):
    """Point ids of the quad at (i,j,k) on the +axis side of the corner
    lattice, ordered so that the right-hand normal points along +axis."""
    if axis == 0:   # x = i plane: (j,k) (j+1,k) (j+1,k+1) (j,k+1)
        return np.stack([pid(i, j, k), pid(i, j + 1, k),
                         pid(i, j + 1, k + 1), pid(i, j, k + 1)], axis=-1)
    if axis == 1:   # y = j plane: (k) along z first then x
        return np.stack([pid(i, j, k), pid(i, j, k + 1),
                         pid(i + 1, j, k + 1), pid(i + 1, j, k)], axis=-1)
    return np.stack([pid(i, j, k), pid(i + 1, j, k),
                     pid(i + 1, j + 1, k), pid(i, j + 1, k)], axis=-1)


def hex_mesh(nx: int, ny: int, nz: int,
             lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
             jitter: float = 0.0, seed: int = DEFAULT_SEED,
             ref_face: bool = True) -> Mesh:
    """Structured box of nx*ny*nz hexahedra in SEMO's unstructured layout.

    Patches, in face order: xmin, xmax, ymin, ymax, [pref], zmin, zmax.
    ``pref`` is the single z-min face of cell 0; it carries the pressure
    reference value that makes the closed-cavity pressure system regular.
    """
    nx, ny, nz = int(nx), int(ny), int(nz)
    px, py, pz = nx + 1, ny + 1, nz + 1
    h = [(hi[a] - lo[a]) / n for a, n in enumerate((nx, ny, nz))]

    def pid(i, j, k):
        return (np.asarray(i) + px * (np.asarray(j) + py * np.asarray(k))).astype(np.int64)

    kk, jj, ii = np.meshgrid(np.arange(pz), np.arange(py), np.arange(px), indexing="ij")
    ii = ii.ravel(); jj = jj.ravel(); kk = kk.ravel()
    pts = np.empty((px * py * pz, 3), dtype=np.float64)
    pts[:, 0] = lo[0] + ii * h[0]
    pts[:, 1] = lo[1] + jj * h[1]
    pts[:, 2] = lo[2] + kk * h[2]
    # exact boundary coordinates (avoid lo + n*h rounding on the far wall)
    pts[ii == nx, 0] = hi[0]; pts[jj == ny, 1] = hi[1]; pts[kk == nz, 2] = hi[2]
    if jitter > 0.0:
        interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz))
        vid = np.nonzero(interior)[0].astype(np.uint64)
        for a in range(3):
            u = unit_uniform(seed, np.uint64(3) * vid + np.uint64(a))
            pts[vid.astype(np.int64), a] += (2.0 * u - 1.0) * jitter * h[a]

    # cells, VTK hexahedron point order
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ci = ci.ravel(); cj = cj.ravel(); ck = ck.ravel()
    ncell = nx * ny * nz
    cell_pts = np.stack([pid(ci, cj, ck), pid(ci + 1, cj, ck), pid(ci + 1, cj + 1, ck),
                         pid(ci, cj + 1, ck), pid(ci, cj, ck + 1), pid(ci + 1, cj, ck + 1),
                         pid(ci + 1, cj + 1, ck + 1), pid(ci, cj + 1, ck + 1)], axis=-1)
    cid = np.arange(ncell, dtype=np.int64)

    # internal faces: per owner in the order +x, +y, +z (neighbour ascending)
    slots_own = np.repeat(cid, 3)
    slots_dir = np.tile(np.arange(3), ncell)
    ci3 = np.repeat(ci, 3); cj3 = np.repeat(cj, 3); ck3 = np.repeat(ck, 3)
    valid = np.where(slots_dir == 0, ci3 < nx - 1,
                     np.where(slots_dir == 1, cj3 < ny - 1, ck3 < nz - 1))
    own = slots_own[valid]; d = slots_dir[valid]
    fi = ci3[valid]; fj = cj3[valid]; fk = ck3[valid]
    nbr = own + np.where(d == 0, 1, np.where(d == 1, nx, nx * ny))
    fpts = np.empty((len(own), 4), dtype=np.int64)
    for a in range(3):
        m = d == a
        oi, oj, ok = fi[m] + (a == 0), fj[m] + (a == 1), fk[m] + (a == 2)
        fpts[m] = _quad_faces_plus(pid, a, oi, oj, ok)

    bowner = []; bpts = []; names = []; starts = []; counts = []
    nf = len(own)

    def add_patch(name, cells, axis, plus):
        nonlocal nf
        cells = np.asarray(cells, dtype=np.int64)
        i0 = ci[cells] + (plus and axis == 0)
        j0 = cj[cells] + (plus and axis == 1)
        k0 = ck[cells] + (plus and axis == 2)
        q = _quad_faces_plus(pid, axis, i0, j0, k0)
        if not plus:
            q = q[:, ::-1]
        names.append(name); starts.append(nf); counts.append(len(cells))
        bowner.append(cells); bpts.append(q); nf += len(cells)

    add_patch("xmin", cid[ci == 0], 0, False)
    add_patch("xmax", cid[ci == nx - 1], 0, True)
    add_patch("ymin", cid[cj == 0], 1, False)
    add_patch("ymax", cid[cj == ny - 1], 1, True)
    zmin_cells = cid[ck == 0]
    if ref_face:
        add_patch("pref", zmin_cells[:1], 2, False)
        add_patch("zmin", zmin_cells[1:], 2, False)
    else:
        add_patch("zmin", zmin_cells, 2, False)
    add_patch("zmax", cid[ck == nz - 1], 2, True)

    owner = np.concatenate([own] + bowner).astype(np.int32)
    face_pts = np.concatenate([fpts] + bpts, axis=0).astype(np.int32)
    nfaces = len(owner)
    return Mesh(points=pts, owner=owner, neighbour=nbr.astype(np.int32),
                face_ptr=(4 * np.arange(nfaces + 1)).astype(np.int32),
                face_pts=face_pts.ravel(),
                cell_ptr=(8 * np.arange(ncell + 1)).astype(np.int32),
                cell_pts=cell_pts.astype(np.int32).ravel(),
                patch_names=names, patch_start=np.array(starts, dtype=np.int32),
                patch_nfaces=np.array(counts, dtype=np.int32), shape=(nx, ny, nz))


def cavity_mesh(n: int, jitter: float = 0.0, seed: int = DEFAULT_SEED) -> Mesh:
    """Unit-cube lid-driven cavity mesh (n^3 hexes) with the pref patch."""
    return hex_mesh(n, n, n, jitter=jitter, seed=seed, ref_face=True)


def _box_points(nx, ny, nz, lo, hi, jitter, seed):
    """Lattice points of a box (x fastest) with the seeded interior jitter of
    ``hex_mesh``; returns the points and their integer lattice coordinates."""
    px, py, pz = nx + 1, ny + 1, nz + 1
    h = [(hi[a] - lo[a]) / n for a, n in enumerate((nx, ny, nz))]
    kk, jj, ii = np.meshgrid(np.arange(pz), np.arange(py), np.arange(px), indexing="ij")
    ii = ii.ravel(); jj = jj.ravel(); kk = kk.ravel()
    pts = np.empty((px * py * pz, 3), dtype=np.float64)
    pts[:, 0] = lo[0] + ii * h[0]
    pts[:, 1] = lo[1] + jj * h[1]
    pts[:, 2] = lo[2] + kk * h[2]
    pts[ii == nx, 0] = hi[0]; pts[jj == ny, 1] = hi[1]; pts[kk == nz, 2] = hi[2]
    if jitter > 0.0:
        interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz))
        vid = np.nonzero(interior)[0].astype(np.uint64)
        for a in range(3):
            u = unit_uniform(seed, np.uint64(3) * vid + np.uint64(a))
            pts[vid.astype(np.int64), a] += (2.0 * u - 1.0) * jitter * h[a]
    return pts, np.stack([ii, jj, kk], axis=1)


# Kuhn split of a cube into 6 tetrahedra along its (0,0,0)-(1,1,1) diagonal:
# for every axis order (a, b, c) the path 0 -> e_a -> e_a + e_b -> (1,1,1).
# Corner ids are bit masks x + 2y + 4z; all cubes use the same diagonal, so
# the split is conforming across cube faces.
_KUHN = [(0, 1 << a, (1 << a) | (1 << b), 7)
         for a, b in ((0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1))]


def tet_mesh(nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
             jitter: float = 0.0, seed: int = DEFAULT_SEED) -> Mesh:
    """Box of nx*ny*nz cubes, each split into 6 tetrahedra (Kuhn split), in
    SEMO's face-based layout: internal faces sorted by (owner, neighbour)
    with owner < neighbour, then the boundary patches xmin, xmax, ymin, ymax,
    zmin, zmax, each sorted by owner.  Face points are ordered so the
    right-hand normal points out of the owner.  Interior lattice points get
    the seeded jitter of ``hex_mesh``."""
    nx, ny, nz = int(nx), int(ny), int(nz)
    px, py = nx + 1, ny + 1
    pts, ijk = _box_points(nx, ny, nz, lo, hi, jitter, seed)
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ci = ci.ravel(); cj = cj.ravel(); ck = ck.ravel()
    ncube = nx * ny * nz

    def corner(b):
        return (ci + (b & 1)) + px * ((cj + ((b >> 1) & 1)) + py * (ck + ((b >> 2) & 1)))

    corners = np.stack([corner(b) for b in range(8)], axis=1).astype(np.int64)
    tets = np.stack([corners[:, list(t)] for t in _KUHN], axis=1).reshape(-1, 4)   # cube-major
    ncell = len(tets)
    # right-handed (VTK) point order: positive signed volume
    x = pts[tets]
    det = np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0])
    neg = det < 0
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    # faces of every tet, matched by their sorted point triples
    tri = np.concatenate([tets[:, [0, 1, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 3]],
                          tets[:, [1, 2, 3]]], axis=0)
    cell_of = np.tile(np.arange(ncell, dtype=np.int64), 4)
    key = np.sort(tri, axis=1)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    first = np.nonzero(same)[0]                       # pairs (first, first+1) in sorted order
    paired = np.zeros(len(ks), dtype=bool); paired[first] = True; paired[first + 1] = True
    a_idx = order[first]; b_idx = order[first + 1]
    ca, cb = cell_of[a_idx], cell_of[b_idx]
    own = np.minimum(ca, cb); nbr = np.maximum(ca, cb)
    itri = np.where((ca <= cb)[:, None], tri[a_idx], tri[b_idx])
    bsel = order[~paired]
    bown = cell_of[bsel]; btri = tri[bsel]
    cen = pts[tets].mean(axis=1)

    def orient(t, owner_cells):
        q = pts[t]
        nrm = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])
        out = np.einsum("ij,ij->i", nrm, q.mean(axis=1) - cen[owner_cells]) > 0
        t = t.copy()
        t[~out] = t[~out][:, [0, 2, 1]]
        return t

    itri = orient(itri, own)
    o = np.lexsort((nbr, own))
    own, nbr, itri = own[o], nbr[o], itri[o]
    btri = orient(btri, bown)
    bijk = ijk[btri]                                   # (nb, 3 points, 3 axes)
    names = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    lim = [0, nx, 0, ny, 0, nz]
    bowner = []; bpts = []; starts = []; counts = []
    nf = len(own)
    taken = np.zeros(len(bown), dtype=bool)
    for pi, nm in enumerate(names):
        ax = pi // 2
        m = np.all(bijk[:, :, ax] == lim[pi], axis=1) & ~taken
        taken |= m
        sel = np.nonzero(m)[0]
        sel = sel[np.argsort(bown[sel], kind="stable")]
        starts.append(nf); counts.append(len(sel)); nf += len(sel)
        bowner.append(bown[sel]); bpts.append(btri[sel])
    assert taken.all(), "boundary face off the box walls"
    owner = np.concatenate([own] + bowner).astype(np.int32)
    face_pts = np.concatenate([itri] + bpts, axis=0).astype(np.int32)
    nfaces = len(owner)
    return Mesh(points=pts, owner=owner, neighbour=nbr.astype(np.int32),
                face_ptr=(3 * np.arange(nfaces + 1)).astype(np.int32),
                face_pts=face_pts.ravel(),
                cell_ptr=(4 * np.arange(ncell + 1)).astype(np.int32),
                cell_pts=tets.astype(np.int32).ravel(),
                patch_names=names, patch_start=np.array(starts, dtype=np.int32),
                patch_nfaces=np.array(counts, dtype=np.int32), shape=(nx, ny, nz))


def rotated(mesh: Mesh, seed: int, shift=(0.0, 0.0, 0.0)) -> Mesh:
    """The mesh turned by a seeded proper rotation about the origin, then
    shifted: every face normal of a box mesh gets three non-zero components,
    and the topology, point orders and patches are unchanged (a proper
    rotation keeps every face's right-hand normal pointing owner ->
    neighbour).  The rotation is the Q factor of a seeded Gaussian matrix with
    positive diagonal of R and det +1."""
    a = np.random.default_rng(seed).standard_normal((3, 3))
    q, r = np.linalg.qr(a)
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0.0:
        q[:, 0] = -q[:, 0]
    pts = np.asarray(mesh.points, dtype=np.float64) @ q.T + np.asarray(shift, dtype=np.float64)
    return Mesh(points=pts, owner=mesh.owner, neighbour=mesh.neighbour, face_ptr=mesh.face_ptr,
                face_pts=mesh.face_pts, cell_ptr=mesh.cell_ptr, cell_pts=mesh.cell_pts,
                patch_names=list(mesh.patch_names), patch_start=mesh.patch_start,
                patch_nfaces=mesh.patch_nfaces, shape=mesh.shape)


def prism_stack_mesh(sides: int = 20, layers: int = 3, radius: float = 1.0, height: float = 0.3,
                     jitter: float = 0.0, seed: int = DEFAULT_SEED) -> Mesh:
    """`layers` prisms over a regular `sides`-gon stacked along z: polyhedral
    cells whose caps (the internal faces between layers and the patches
    "bottom" / "top") have `sides` points, side walls of quads (patch
    "side").  Every point moves by a seeded U(-jitter, jitter) times the
    polygon's edge length in each coordinate, so faces are not planar.
    Faces point owner -> neighbour (internal: +z), boundary faces outward."""
    rng = np.random.default_rng(seed)
    th = 2.0 * np.pi * np.arange(sides) / sides
    edge = 2.0 * radius * np.sin(np.pi / sides)
    pts = np.array([[radius * np.cos(t), radius * np.sin(t), l * height]
                    for l in range(layers + 1) for t in th])
    pts += jitter * edge * rng.uniform(-1.0, 1.0, pts.shape)
    ring = lambda l: [l * sides + i for i in range(sides)]     # noqa: E731  counter-clockwise seen from +z
    faces, owner, nbr = [], [], []
    for l in range(1, layers):                     # internal caps, owner below
        faces.append(ring(l)); owner.append(l - 1); nbr.append(l)
    names, starts, counts = ["bottom", "top", "side"], [], []
    starts.append(len(faces)); faces.append(ring(0)[::-1]); owner.append(0); counts.append(1)
    starts.append(len(faces)); faces.append(ring(layers)); owner.append(layers - 1); counts.append(1)
    starts.append(len(faces))
    for l in range(layers):
        lo, hi = ring(l), ring(l + 1)
        for i in range(sides):
            j = (i + 1) % sides
            faces.append([lo[i], lo[j], hi[j], hi[i]]); owner.append(l)
    counts.append(len(faces) - starts[2])
    face_ptr = np.concatenate([[0], np.cumsum([len(f) for f in faces])]).astype(np.int32)
    cells = [ring(l) + ring(l + 1) for l in range(layers)]
    cell_ptr = np.concatenate([[0], np.cumsum([len(c) for c in cells])]).astype(np.int32)
    return Mesh(points=pts, owner=np.array(owner, dtype=np.int32), neighbour=np.array(nbr, dtype=np.int32),
                face_ptr=face_ptr, face_pts=np.array([p for f in faces for p in f], dtype=np.int32),
                cell_ptr=cell_ptr, cell_pts=np.array([p for c in cells for p in c], dtype=np.int32),
                patch_names=names, patch_start=np.array(starts, dtype=np.int32),
                patch_nfaces=np.array(counts, dtype=np.int32))
