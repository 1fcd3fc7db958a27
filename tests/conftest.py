import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsemo_gpu.so)")


def ensure_oracle():
    so = os.path.join(ROOT, "oracle", "liborc.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile")],
                       check=True, cwd=ROOT)
    return so


@pytest.fixture(scope="session")
def oracle_lib():
    ensure_oracle()
    from oracle import oracle
    return oracle


def load_golden(name):
    """Fixture arrays and the mesh they were computed on."""
    from semo_mesh_amd.mesh import Mesh
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    d = {k: z[k] for k in z.files}
    mesh = Mesh(points=d["mesh_points"], owner=d["mesh_owner"], neighbour=d["mesh_neighbour"],
                face_ptr=d["mesh_face_ptr"], face_pts=d["mesh_face_pts"],
                cell_ptr=d["mesh_cell_ptr"], cell_pts=d["mesh_cell_pts"],
                patch_names=[str(s) for s in d["mesh_patch_names"]],
                patch_start=d["mesh_patch_start"], patch_nfaces=d["mesh_patch_nfaces"])
    return d, mesh


def golden_case(d):
    from semo_mesh_amd.cases import cavity_case
    c = cavity_case(dt=float(d["case_dt"]), lid=tuple(d["case_lid"]))
    c.mu = tuple(d["case_mu"]); c.rho = tuple(d["case_rho"])
    return c


GOLDEN_CASES = ["cavity4_jitter_random", "box654_jitter_2x2", "cavity8_plain", "orifice_lid"]
