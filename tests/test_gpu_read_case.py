"""Cases read from SEMO's own files, run on the GPU against the oracle.

* The state SEMO_Mesh_Save::vtu wrote (tests/golden/vtu_zlib_tet221.vtu, the
  zlib encoding): mesh, P, U, V, W and the volume fraction as the reader
  gives them, two two-phase time steps (dam-break conditions).
* A jittered 6^3 cube whose starting fields come from the initial-field
  region initialisers (the spec of tests/golden/initial_cube6.npz whose
  output SEMO's initial utility is pinned to), written as SEMO's binary VTU
  and read back: three SIMPLE iterations.
Every field within relative L2 1e-10 (FP64) / 1e-6 (mixed) of the oracle.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden
from semo_mesh_amd.cases import cavity_case, dam_break_case
from semo_mesh_amd.vtu import read_vtu, semo_fields, write_vtu

pytestmark = pytest.mark.gpu


def _compare(g, o, fields, tol):
    vel = np.sqrt(sum(np.linalg.norm(o.get(f)) ** 2 for f in "UVW"))
    for f in fields:
        a, b = g.get_field(f), o.get(f)
        scale = np.linalg.norm(b) if f in ("P", "VF") else vel
        err = np.linalg.norm(a - b) / scale
        print(f, err)
        assert err <= tol, (f, err)


def _solver(mesh, case, precision):
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    g.upload_mesh(mesh, reorder=True)
    g.set_case(case)
    g.set_solver(precision, tol_p=1e-13 if precision == "fp64" else 1e-10, tol_u=1e-13, maxit_p=3000,
                 maxit_u=3000)
    return g


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_semo_saved_state_vof_steps(oracle_lib, precision, tol):
    mesh, f, _ = read_vtu(os.path.join(GOLDEN, "vtu_zlib_tet221.vtu"))
    init = {"P": f["pressure"], "U": f["velocity"][:, 0].copy(), "V": f["velocity"][:, 1].copy(),
            "W": f["velocity"][:, 2].copy(), "VF": f["volume-fraction-water"]}
    case = dam_break_case(dt=2e-3, n_vof=2, urf_vf=0.9)
    o = oracle_lib.Oracle(mesh, case)
    o.set_solver(1e-14, 10000)
    g = _solver(mesh, case, precision)
    try:
        for k, v in init.items():
            o.set(k, v)
            g.set_field(k, v)
        for _ in range(2):
            o.step(2)
            g.incompressiblePressureBased(2)
        _compare(g, o, ("P", "U", "V", "W", "VF"), tol)
    finally:
        g.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_initialised_cube_simple_iterations(oracle_lib, tmp_path, precision, tol):
    from semo_mesh_amd.initial import initial_fields
    d, m = load_golden("initial_cube6")
    const = os.path.join(tmp_path, "constant")
    os.makedirs(const)
    for fname, body in zip(("pressure", "velocities", "temperature", "water"), d["box_cyl_box_sphere_spec"]):
        with open(os.path.join(const, fname), "w") as fh:
            fh.write("initialField\n{\n" + "".join(f"\t{ln}\n" for ln in str(body).split("\n")) + "}\n")
    ini = initial_fields(str(tmp_path), m)
    p = os.path.join(tmp_path, "plot.0.vtu")
    write_vtu(p, m, semo_fields(ini["P"], ini["U"], ini["V"], ini["W"], VF=ini["VF"], T=ini["T"]), fmt="binary")
    mesh, f, _ = read_vtu(p)
    case = cavity_case()
    case.mu = (0.01, 0.01)
    o = oracle_lib.Oracle(mesh, case)
    o.set_solver(1e-14, 10000)
    g = _solver(mesh, case, precision)
    try:
        for k, v in (("P", f["pressure"]), ("U", f["velocity"][:, 0].copy()), ("V", f["velocity"][:, 1].copy()),
                     ("W", f["velocity"][:, 2].copy())):
            o.set(k, v)
            g.set_field(k, v)
        for _ in range(3):
            o.step(1)
            g.incompressiblePressureBased(1)
        _compare(g, o, ("P", "U", "V", "W"), tol)
    finally:
        g.close()
