"""semo_least_square_noalloc (the LSQ gradient on a static device scratch,
no allocation) against semo_least_square on the same inputs: bit-identical,
on two calls in a row on one GPU, the second on a smaller mesh, and a mesh
larger than the scratch is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from semo_mesh_amd.mesh import cavity_mesh, hex_mesh, tet_mesh

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers through the HIP runtime libsemo_gpu.so itself uses
    (torch carries a runtime of its own, which does not share its device
    state with this one)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p.value

    def get(self, ptr, n):
        out = np.empty(n, dtype=np.float64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _check(g, mesh, seed):
    rng = np.random.default_rng(seed)
    n, nif, nf = mesh.n_cells, mesh.n_internal_faces, mesh.n_faces
    phi = rng.standard_normal(n)
    fr = rng.standard_normal(nf)
    ref = g.calcLeastSquare(phi, fr)                 # host arrays, caller numbering
    perm, _ = g.permutation()                        # device cell k = caller cell perm[k]
    fmap, _ = g.face_map()                           # device face k = caller face fmap[k]
    dv = _Dev()
    d_phi = dv.put(phi[perm])
    d_fr = dv.put(fr[fmap[nif:]])
    d_g = dv.put(np.full(3 * n, np.nan))
    g.least_square_noalloc(d_phi, d_fr, d_g)
    dev = dv.get(d_g, 3 * n).reshape(n, 3)
    dv.free()
    got = np.empty_like(dev)
    got[perm] = dev
    assert np.array_equal(got, ref)


def test_noalloc_lsq_matches_allocating_entry_point():
    from semo_mesh_amd.solver import GpuSolver
    g = GpuSolver(0)
    m1 = cavity_mesh(16, jitter=0.1, seed=51)
    g.upload_mesh(m1, reorder=True)
    _check(g, m1, 1)
    m2 = tet_mesh(5, 4, 3, jitter=0.1, seed=52)      # second call: fewer cells, other stencils
    assert m2.n_cells < m1.n_cells
    g.upload_mesh(m2, reorder=True)
    _check(g, m2, 2)
    g.close()


def test_noalloc_lsq_refuses_meshes_beyond_its_scratch():
    from semo_mesh_amd._lib import SemoError
    from semo_mesh_amd.solver import GpuSolver
    m = hex_mesh(129, 128, 128, ref_face=False)     # 2,113,536 cells > 2,097,152
    g = GpuSolver(0)
    g.upload_mesh(m, reorder=False)
    dv = _Dev()
    d = dv.put(np.zeros(3 * m.n_cells))
    with pytest.raises(SemoError, match="SEMO_EINVAL"):
        g.least_square_noalloc(d, d, d)
    dv.free()
    g.close()
