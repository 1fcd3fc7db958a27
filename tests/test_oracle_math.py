"""Closed-form and structural checks on the CPU oracle, independent of the
reference fixtures: a dropped term, a wrong sign or a transposed index in the
restatement breaks one of these."""
import numpy as np
import pytest

from semo_mesh_amd.cases import cavity_case
from semo_mesh_amd.mesh import cavity_mesh, hex_mesh


def test_uniform_hex_geometry_closed_form(oracle_lib):
    n = 5
    h = 1.0 / n
    m = cavity_mesh(n)
    o = oracle_lib.Oracle(m, cavity_case())
    assert np.allclose(o.get("face_area"), h * h, rtol=1e-13)
    assert np.allclose(o.get("cell_volume"), h ** 3, rtol=1e-12)
    cells = m.cell_pts.reshape(-1, 8)
    centre = m.points[cells].mean(axis=1)
    assert np.allclose(o.get("cell_xyz").reshape(-1, 3), centre, atol=1e-14)
    assert np.all(o.get("face_wC") == 0.5)
    assert np.allclose(o.get("face_alphaF"), 1.0, rtol=1e-13)
    assert np.allclose(o.get("face_skew"), 0.0, atol=1e-15)
    assert np.allclose(o.get("face_magPN"), h, rtol=1e-12)      # boundary: 2 * h/2
    nrm = o.get("face_normal").reshape(-1, 3)
    nif = m.n_internal_faces
    d = o.get("cell_xyz").reshape(-1, 3)
    e = d[m.neighbour] - d[m.owner[:nif]]
    assert np.all(np.einsum("ij,ij->i", nrm[:nif], e) > 0)         # owner -> neighbour


def test_jittered_volumes_sum_to_domain(oracle_lib):
    m = hex_mesh(6, 5, 4, jitter=0.1, seed=3)
    o = oracle_lib.Oracle(m, cavity_case())
    assert abs(o.get("cell_volume").sum() - 1.0) < 1e-13
    # divergence theorem: sum over a closed cell of n*A is zero
    nrm = o.get("face_normal").reshape(-1, 3) * o.get("face_area")[:, None]
    acc = np.zeros((m.n_cells, 3))
    np.add.at(acc, m.owner, nrm)
    np.add.at(acc, m.neighbour, -nrm[: m.n_internal_faces])
    assert np.abs(acc).max() < 1e-14


def test_lsq_gradient_exact_on_linear_field(oracle_lib):
    m = hex_mesh(7, 6, 5, jitter=0.1, seed=5)
    o = oracle_lib.Oracle(m, cavity_case())
    a = np.array([0.3, -1.7, 2.2])
    c = o.get("cell_xyz").reshape(-1, 3)
    f = o.get("face_xyz").reshape(-1, 3)
    phi = 4.0 + c @ a
    face_r = 4.0 + f @ a
    g = o.lsq(phi, face_r)
    assert np.allclose(g, a[None, :], atol=1e-12)


def test_minmax_brute_force(oracle_lib):
    m = hex_mesh(4, 3, 5, jitter=0.1, seed=9)
    o = oracle_lib.Oracle(m, cavity_case())
    phi = np.random.default_rng(0).normal(size=m.n_cells)
    mx, mn = o.minmax(phi)
    # brute force: cells sharing at least one point
    cp = m.cell_pts.reshape(-1, 8)
    for i in range(m.n_cells):
        nb = [j for j in range(m.n_cells) if np.intersect1d(cp[i], cp[j]).size]
        assert mx[i] == phi[nb].max() and mn[i] == phi[nb].min()


def _bandwidth(ptr, adj, inv):
    bw = 0
    for i in range(len(ptr) - 1):
        for k in range(ptr[i], ptr[i + 1]):
            bw = max(bw, abs(int(inv[i]) - int(inv[adj[k]])))
    return bw


def test_rcm_path_graph_and_bandwidth(oracle_lib):
    # a path 0-1-...-n-1 numbered at random: RCM must give bandwidth 1
    rng = np.random.default_rng(1)
    n = 40
    lab = rng.permutation(n)
    nbrs = [[] for _ in range(n)]
    for k in range(n - 1):
        a, b = lab[k], lab[k + 1]
        nbrs[a].append(b); nbrs[b].append(a)
    ptr = np.zeros(n + 1, dtype=np.int32)
    ptr[1:] = np.cumsum([len(x) for x in nbrs])
    adj = np.array([y for x in nbrs for y in sorted(x)], dtype=np.int32)
    perm, inv = oracle_lib.genrcm(ptr, adj)
    assert sorted(perm) == list(range(n))
    assert np.array_equal(inv[perm], np.arange(n))
    assert _bandwidth(ptr, adj, inv) == 1
    # the reversal: the pseudo-peripheral start ends last
    assert perm[-1] in (lab[0], lab[-1])


def test_rcm_reduces_bandwidth_of_shuffled_mesh(oracle_lib):
    m = hex_mesh(6, 6, 6)
    o = oracle_lib.Oracle(m, cavity_case())
    ptr, adj = o.cell_graph()
    n = m.n_cells
    sh = np.random.default_rng(2).permutation(n).astype(np.int32)   # new label of old cell
    inv_sh = np.argsort(sh).astype(np.int32)
    # relabelled graph
    ptr2 = np.zeros(n + 1, dtype=np.int32); adj2 = []
    for new in range(n):
        old = inv_sh[new]
        nb = sorted(sh[adj[ptr[old]:ptr[old + 1]]])
        adj2 += nb; ptr2[new + 1] = len(adj2)
    adj2 = np.array(adj2, dtype=np.int32)
    perm, inv = oracle_lib.genrcm(ptr2, adj2)
    assert _bandwidth(ptr2, adj2, inv) <= 2 * 36 and _bandwidth(ptr2, adj2, np.arange(n)) > 100


def test_pressure_matrix_symmetric_laplacian(oracle_lib):
    m = hex_mesh(5, 4, 6, jitter=0.1, seed=4)
    o = oracle_lib.Oracle(m, cavity_case())
    o.eos(); o.recon(); o.assemble_pressure()
    r, c, v = o.coo()
    import scipy.sparse as sp
    A = sp.coo_matrix((v, (r, c)), shape=(m.n_cells,) * 2).tocsr()
    assert abs(A - A.T).max() == 0.0
    rs = np.asarray(A.sum(axis=1)).ravel()
    # zero row sums except at the pref cell (fixedValue p adds to its diagonal)
    pref = m.owner[m.patch_start[m.patch("pref")]]
    assert np.all(np.abs(np.delete(rs, pref)) < 1e-12 * np.abs(v).max())
    assert rs[pref] > 0


def test_inlet_profile_shape_and_bounds():
    """cases.inlet_profile: [3][n_boundary_faces], the perturbation (amplitude
    2 %) only on the named patch, (u_in, 0, 0) elsewhere."""
    from semo_mesh_amd.cases import inlet_profile
    m = hex_mesh(8, 4, 4, lo=(1.0, 0.0, 0.0), hi=(5.0, 1.0, 1.0), jitter=0.1, seed=3, ref_face=False)
    u = inlet_profile(m, "xmin", 2.0, amplitude=0.02, seed=7)
    nb = m.n_faces - m.n_internal_faces
    assert u.shape == (3, nb)
    b = m.patch("xmin")
    s = int(m.patch_start[b]) - m.n_internal_faces
    sl = slice(s, s + int(m.patch_nfaces[b]))
    inlet = u[0, sl]
    assert np.all(np.abs(inlet / 2.0 - 1.0) <= 0.02 + 1e-15) and np.ptp(inlet) > 0.0
    assert np.all(u[1:] == 0.0)
    rest = np.delete(u[0], np.arange(nb)[sl])
    assert np.all(rest == 2.0)
    assert np.array_equal(u, inlet_profile(m, "xmin", 2.0, amplitude=0.02, seed=7))


def test_oracle_uniform_face_velocity_is_the_patch_value(oracle_lib):
    """The oracle's per-face fixedValue-U values, set to the patch value on
    every boundary face, reproduce the uniform inlet bit for bit; a profile
    changes the solution."""
    from semo_mesh_amd.cases import channel_case, inlet_profile
    m = hex_mesh(8, 3, 3, lo=(1.0, 0.0, 0.0), hi=(5.0, 1.0, 1.0), jitter=0.1, seed=5, ref_face=False)
    case = channel_case(1.0, dt=0.02, rho=1.0, mu=0.01)
    nb = m.n_faces - m.n_internal_faces
    uni = np.zeros((3, nb)); uni[0] = 1.0
    out = []
    for ufix in (None, uni, inlet_profile(m, "xmin", 1.0)):
        o = oracle_lib.Oracle(m, case)
        o.set_solver(tol=1e-14)
        o.set("U", np.full(m.n_cells, 1.0))
        if ufix is not None:
            o.set("ufix", ufix.ravel())
        o.step(correctors=2)
        out.append(o.get("U"))
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[0], out[2])


def test_gauss_green_exact_for_linear_fields(oracle_lib):
    """On an unjittered box (planar faces, wC = 0.5, no skewness) the
    Green-Gauss gradient of a linear field, with zeroGradient walls
    extrapolating with the exact gradient, is that gradient (divergence
    theorem); a fixedValue wall whose value is the field's value there keeps
    it exact."""
    from dataclasses import replace
    from semo_mesh_amd.cases import BC_FIXED_VALUE, BC_ZERO_GRADIENT, cavity_case
    from semo_mesh_amd.mesh import hex_mesh
    m = hex_mesh(5, 4, 3, hi=(1.0, 0.8, 0.6), ref_face=False)
    case = cavity_case()
    case.bcs = {nm: replace(case.bcs.get(nm, case.bcs["xmin"]), p_type=BC_ZERO_GRADIENT)
                for nm in m.patch_names}
    # phi = 2 + 3 (x - 0) - y + 0.5 z: zero on no wall; put xmin (x = 0) at
    # its value along a field constant in y, z by using the x-only part
    case.bcs["xmin"] = replace(case.bcs["xmin"], p_type=BC_FIXED_VALUE, p_value=2.0)
    o = oracle_lib.Oracle(m, case)
    x = o.get("cell_xyz").reshape(-1, 3)
    a = np.array([3.0, 0.0, 0.0])
    phi = 2.0 + x @ a
    grad = np.tile(a, (m.n_cells, 1))
    g = o.gauss_green("P", phi, np.zeros(m.n_faces), grad)
    assert np.abs(g - a).max() <= 1e-12 * 3.0
    # a wrong boundary value breaks it (the check is sensitive to the wall)
    case.bcs["xmin"] = replace(case.bcs["xmin"], p_value=2.5)
    o2 = oracle_lib.Oracle(m, case)
    g2 = o2.gauss_green("P", phi, np.zeros(m.n_faces), grad)
    assert np.abs(g2 - a).max() > 1e-3
