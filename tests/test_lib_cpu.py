"""CPU-side checks of the product library: it loads, exports every symbol its
header declares, and its host-only entry point (GENRCM) reproduces the
reference permutation bit for bit."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN_CASES, ROOT, load_golden


def _lib_or_skip():
    from semo_mesh_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsemo_gpu.so not built")
    return _lib


def test_header_symbols_exported():
    _lib = _lib_or_skip()
    hdr = open(os.path.join(ROOT, "include", "semo_gpu.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(semo_\w+)\s*\(", hdr, re.M))
    assert declared == set(_lib.EXPORTED)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_library_rcm_matches_reference(name, oracle_lib):
    _lib_or_skip()
    from semo_mesh_amd.solver import rcm
    d, mesh = load_golden(name)
    o = oracle_lib.Oracle(mesh)
    ptr, adj = o.cell_graph()
    perm, inv = rcm(ptr, adj)
    assert np.array_equal(perm, d["rcm_perm"]) and np.array_equal(inv, d["rcm_inv"])


def test_library_rcm_matches_oracle_on_random_graphs(oracle_lib):
    _lib_or_skip()
    from semo_mesh_amd.solver import rcm
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(1, 60))
        m = int(rng.integers(0, 3 * n))
        e = rng.integers(0, n, size=(m, 2))
        e = e[e[:, 0] != e[:, 1]]
        nb = [set() for _ in range(n)]
        for a, b in e:
            nb[a].add(int(b)); nb[b].add(int(a))
        order = [list(rng.permutation(sorted(s))) for s in nb]     # arbitrary adjacency order
        ptr = np.zeros(n + 1, dtype=np.int32); ptr[1:] = np.cumsum([len(x) for x in order])
        adj = np.array([y for x in order for y in x], dtype=np.int32)
        p1, i1 = rcm(ptr, adj)
        p0, i0 = oracle_lib.genrcm(ptr, adj)
        assert np.array_equal(p1, p0) and np.array_equal(i1, i0), trial


def test_rcm_rejects_bad_graph():
    _lib = _lib_or_skip()
    from semo_mesh_amd.solver import rcm
    with pytest.raises(_lib.SemoError):
        rcm(np.array([0, 1], dtype=np.int32), np.array([5], dtype=np.int32))
