"""Per-face fixedValue-U values (SEMO_F_FACE_UFIX): an inlet profile.

SEMO gives fixedValue-U boundaries one value per patch and shapes inlets
through UDFs (reconIncom.cpp:707-727).  The per-face override is checked:
  * against the oracle carrying the same override (FP64, 3 SIMPLE iterations
    of a jittered channel with a 2 % sinusoidal inlet perturbation), with
    and without RCM renumbering: fields to 1e-10;
  * values on faces that are not fixedValue-U are ignored (bit for bit);
  * a profile equal to the patch value everywhere, and clearing the
    profile, reproduce the uniform inlet bit for bit.
"""
import numpy as np
import pytest

from semo_mesh_amd.cases import channel_case
from semo_mesh_amd.mesh import hex_mesh

pytestmark = pytest.mark.gpu

U_IN = 1.0


def _mesh():
    # x in [1, 5]: away from the live orifice UDF near x = 0.0006 k
    return hex_mesh(12, 4, 4, lo=(1.0, 0.0, 0.0), hi=(5.0, 1.0, 1.0), jitter=0.1, seed=41, ref_face=False)


def _case():
    c = channel_case(U_IN, dt=0.02, rho=1.0, mu=0.01)
    c.urf_mom, c.urf_p, c.urf_pu = 0.7, 0.3, 0.3
    return c


def _profile(mesh, fill_others=0.0):
    nb = mesh.n_faces - mesh.n_internal_faces
    u = np.full((3, nb), fill_others)
    b = mesh.patch("xmin")
    s, n = int(mesh.patch_start[b]) - mesh.n_internal_faces, int(mesh.patch_nfaces[b])
    # face centres of the inlet from the point lists
    fxyz = np.array([mesh.points[mesh.face_pts[mesh.face_ptr[f]:mesh.face_ptr[f + 1]]].mean(axis=0)
                     for f in range(s + mesh.n_internal_faces, s + mesh.n_internal_faces + n)])
    ph = np.random.default_rng(43).random(2)
    u[:, s:s + n] = 0.0
    u[0, s:s + n] = U_IN * (1.0 + 0.02 * np.sin(2 * np.pi * (fxyz[:, 1] + ph[0])) * np.sin(2 * np.pi * (fxyz[:, 2] + ph[1])))
    return u


def _gpu(mesh, reorder, prof=None):
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    g.upload_mesh(mesh, reorder=reorder)
    g.set_case(_case())
    g.set_solver("fp64", tol_p=1e-13, tol_u=1e-13, maxit_p=500, maxit_u=500)
    g.set_field("U", np.full(mesh.n_cells, U_IN))
    if prof is not None:
        g.set_boundary_velocity(prof)
    return g


def _fields(g):
    return {f: g.get_field(f) for f in ("P", "U", "V", "W")}


@pytest.mark.parametrize("reorder", [False, True])
def test_inlet_profile_matches_oracle(oracle_lib, reorder):
    mesh = _mesh()
    prof = _profile(mesh)
    g = _gpu(mesh, reorder, prof)
    assert np.array_equal(g.get_field("face_Ufix"), prof)
    g.incompressiblePressureBased(3)
    fg = _fields(g)
    g.close()
    o = oracle_lib.Oracle(mesh, _case())
    o.set_solver(tol=1e-14)
    o.set("U", np.full(mesh.n_cells, U_IN))
    o.set("ufix", prof.ravel())
    o.step(correctors=3)
    for f in fg:
        ref = o.get(f)
        rel = np.linalg.norm(fg[f] - ref) / max(np.linalg.norm(ref), 1e-300)
        print(f, rel)
        assert rel <= 1e-10, f


def test_inlet_profile_ignores_other_faces_and_uniform_is_identity():
    mesh = _mesh()
    a = _gpu(mesh, True, _profile(mesh, 0.0))
    a.incompressiblePressureBased(2)
    b = _gpu(mesh, True, _profile(mesh, 1e3))      # values on walls / outlet: ignored
    b.incompressiblePressureBased(2)
    fa, fb = _fields(a), _fields(b)
    for f in fa:
        assert np.array_equal(fa[f], fb[f]), f
    a.close(); b.close()
    nb = mesh.n_faces - mesh.n_internal_faces
    uni = np.zeros((3, nb)); uni[0] = U_IN
    plain = _gpu(mesh, True)
    plain.incompressiblePressureBased(2)
    same = _gpu(mesh, True, uni)
    same.incompressiblePressureBased(2)
    cleared = _gpu(mesh, True, _profile(mesh))
    cleared.set_boundary_velocity(None)
    cleared.incompressiblePressureBased(2)
    fp, fs, fc = _fields(plain), _fields(same), _fields(cleared)
    for f in fp:
        assert np.array_equal(fp[f], fs[f]), f
        assert np.array_equal(fp[f], fc[f]), f
    plain.close(); same.close(); cleared.close()
