"""Pins of the oracle's CVFE dual (oracle/orc_mesh.c) against closed forms,
invariants and the reference's own Jacobian pattern fixture."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.oracle import OracleMesh


def _meshes(cassion_mesh, slice2d, box3d_small):
    return [("cassion", cassion_mesh), ("slice2d", slice2d.mesh), ("box3d", box3d_small.mesh)]


@pytest.fixture(scope="module")
def built(cassion_mesh, slice2d, box3d_small):
    return [(n, m, OracleMesh(m)) for n, m in _meshes(cassion_mesh, slice2d, box3d_small)]


def test_scv_volumes_are_simplex_fractions(built):
    for name, m, om in built:
        V, scv = om.volumes()
        vol = np.abs(m.element_volumes())
        np.testing.assert_allclose(scv, np.repeat(vol[:, None] / (m.dim + 1), m.dim + 1, 1), rtol=1e-11,
                                   err_msg=name)


def test_cv_volumes_sum_to_domain(built):
    expect = {"slice2d": 50.0, "box3d": 100.0 * 100.0 * 20.0}
    for name, m, om in built:
        V, _ = om.volumes()
        total = expect.get(name, np.abs(m.element_volumes()).sum())
        assert abs(V.sum() - total) <= 1e-10 * total, name


def test_cv_surfaces_are_closed_and_oriented(built):
    for name, m, om in built:
        ints, dbl = om.cvfaces()
        elem, edge, front, back, tag = ints.T
        area, n = dbl[:, 0], dbl[:, 1:4]
        nA = n * area[:, None]
        interior = front >= 0
        # normals of interior faces point back -> front (cvface_inline.h:112-119)
        d = m.xyz[front[interior]] - m.xyz[back[interior]]
        assert np.all(np.einsum("ij,ij->i", nA[interior], d) > 0), name
        # each control volume is closed: outward area vectors sum to zero
        s = np.zeros((m.n_nodes, 3))
        np.add.at(s, back, nA)
        np.add.at(s, front[interior], -nA[interior])
        scale = np.abs(nA).max()
        assert np.abs(s).max() < 1e-12 * scale * 10, (name, np.abs(s).max())
        # interior faces sorted by (edge, element) (mesh.cpp:802)
        key = edge[interior].astype(np.int64) * (m.n_elements + 1) + elem[interior]
        assert np.all(np.diff(key) > 0), name


def test_face_area_vectors_match_gradient_identity(built):
    """S_ab = |T|/(d+1) (grad N_b - grad N_a): the closed form the HIP kernel uses."""
    for name, m, om in built:
        ints, dbl = om.cvfaces()
        g = om.gradients()
        vol = np.abs(m.element_volumes())
        nA = dbl[:, 1:4] * dbl[:, 0:1]
        for f in range(0, om.NFI, max(1, om.NFI // 2000)):
            e, fr, bk = ints[f, 0], ints[f, 2], ints[f, 3]
            lf = int(np.nonzero(m.conn[e] == fr)[0][0])
            lb = int(np.nonzero(m.conn[e] == bk)[0][0])
            S = vol[e] / (m.dim + 1) * (g[e, lf] - g[e, lb])
            assert np.allclose(S, nA[f], rtol=1e-10, atol=1e-12 * np.abs(nA[f]).max()), (name, f, S, nA[f])


def test_gradients_exact_for_linear_fields(built):
    rng = np.random.default_rng(5)
    for name, m, om in built:
        a = rng.normal(size=3)
        if m.dim == 2:
            a[2] = 0
        h = m.xyz @ a + 0.7
        g = om.gradients()
        gh = np.einsum("tkd,tk->td", g, h[m.conn])
        np.testing.assert_allclose(gh, np.broadcast_to(a, gh.shape), atol=1e-9, err_msg=name)


def test_edges_first_insertion_and_sorted(built):
    for name, m, om in built:
        f, b = om.edges()
        lo, hi = np.minimum(f, b), np.maximum(f, b)
        key = lo.astype(np.int64) * m.n_nodes + hi
        assert np.all(np.diff(key) > 0), name
        # front of each edge is the first node of the first element edge listing it
        tab = [[0, 1], [1, 2], [2, 0]] if m.dim == 2 else [[0, 1], [1, 2], [2, 0], [0, 3], [1, 3], [2, 3]]
        seen = {}
        for e in range(m.n_elements):
            for a_, b_ in tab:
                n0, n1 = int(m.conn[e, a_]), int(m.conn[e, b_])
                k = (min(n0, n1), max(n0, n1))
                if k not in seen:
                    seen[k] = n0
        idx = {(int(lo[i]), int(hi[i])): i for i in range(len(lo))}
        for k, n0 in list(seen.items())[:5000]:
            assert f[idx[k]] == n0


def test_jacobian_pattern_matches_reference_fixture(built):
    """meshing/Jpattern.mtx (decomp.c:92): node adjacency plus diagonal of the cassion mesh."""
    name, m, om = built[0]
    rows, cols = [], []
    with open(os.path.join(GOLDEN, "Jpattern.mtx")) as fh:
        header = fh.readline()
        assert "MatrixMarket" in header
        line = fh.readline()
        while line.startswith("%"):
            line = fh.readline()
        n, n2, nnz = (int(v) for v in line.split())
        for _ in range(nnz):
            r, c, _v = fh.readline().split()
            rows.append(int(r) - 1)
            cols.append(int(c) - 1)
    assert n == m.n_nodes and nnz == m.n_nodes + 2 * om.NE
    f, b = om.edges()
    ours = set(zip(range(n), range(n))) | set(zip(f.tolist(), b.tolist())) | set(zip(b.tolist(), f.tolist()))
    assert ours == set(zip(rows, cols))
