"""SEMO's plot.<rank>.vtu, ASCII (semo_mesh_amd.vtu), CPU only.

The reference holds no .vtu case, so the reader is pinned by round trips of
the writer and by the loader's rules (src/load/load.cpp:352-413): every mesh
array and field comes back identical, the oracle's geometry and one SIMPLE
iteration on the read-back mesh are bit-identical to the original's, and a
face's points come from the first cell whose face stream lists it.
"""
import os

import numpy as np

from semo_mesh_amd.cases import cavity_case
from semo_mesh_amd.mesh import hex_mesh, tet_mesh
from semo_mesh_amd.vtu import read_vtu, write_vtu

ARRAYS = ("points", "owner", "neighbour", "face_ptr", "face_pts", "cell_ptr", "cell_pts",
          "patch_start", "patch_nfaces")


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    for k, m in enumerate((hex_mesh(4, 3, 2, jitter=0.1, seed=5), tet_mesh(2, 1, 1, jitter=0.1))):
        p = os.path.join(tmp_path, f"plot.{k}.vtu")
        f = {"pressure": rng.standard_normal(m.n_cells), "velocity": rng.standard_normal((m.n_cells, 3))}
        write_vtu(p, m, f, time=0.125)
        r, g, t = read_vtu(p)
        for a in ARRAYS:
            assert np.array_equal(getattr(r, a), getattr(m, a)), a
        assert r.patch_names == m.patch_names
        assert t == 0.125
        assert np.array_equal(g["pressure"], f["pressure"]) and np.array_equal(g["velocity"], f["velocity"])


def test_oracle_bit_identical_on_read_mesh(tmp_path, oracle_lib):
    m = hex_mesh(5, 4, 3, jitter=0.1, seed=13, ref_face=True)
    p = os.path.join(tmp_path, "plot.0.vtu")
    write_vtu(p, m)
    r, _, _ = read_vtu(p)
    case = cavity_case(dt=0.05)
    case.rho = (1.0, 1.0)
    case.mu = (0.01, 0.01)
    res = []
    for mesh in (m, r):
        o = oracle_lib.Oracle(mesh, case)
        o.set_solver(tol=1e-14)
        o.step(correctors=1)
        res.append({f: o.get(f) for f in ("P", "U", "V", "W", "cell_volume", "face_area")})
    for f in res[0]:
        assert np.array_equal(res[0][f], res[1][f]), f


def test_face_points_from_the_first_cell(tmp_path):
    """An internal face listed with its points reversed in the neighbour's
    stream keeps the owner's order (the owner, lower index, comes first)."""
    m = hex_mesh(2, 1, 1)
    p = os.path.join(tmp_path, "plot.0.vtu")
    write_vtu(p, m)
    text = open(p).read()
    f0 = m.face_pts[m.face_ptr[0]:m.face_ptr[1]].tolist()          # the internal face
    own = "4 " + " ".join(map(str, f0))
    rev = "4 " + " ".join(map(str, f0[::-1]))
    head, tail = text.split('Name="faces" format="ascii">\n', 1)
    stream, rest = tail.split("\n", 1)
    i = stream.rfind(own)                 # the neighbour cell's copy is the last one
    assert i > stream.find(own)
    stream = stream[:i] + rev + stream[i + len(own):]
    open(p, "w").write(head + 'Name="faces" format="ascii">\n' + stream + "\n" + rest)
    r, _, _ = read_vtu(p)
    assert r.face_pts[r.face_ptr[0]:r.face_ptr[1]].tolist() == f0
