"""Reverse Cuthill-McKee renumbering (mesh.cpp:118-209, 1571-1615; Boost's
cuthill_mckee_ordering, which is not available here).

Pins: the oracle's sweep, started at the node scipy's published RCM starts
at, reproduces scipy.sparse.csgraph.reverse_cuthill_mckee node for node (same
rule: unvisited neighbours in increasing degree, ties in adjacency order);
its pseudo-peripheral start (George and Liu, as Boost's) has an eccentricity
no smaller than the lowest-degree node's; the library's fvp_rcm equals the
oracle's order; both are permutations and cut the bandwidth."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import breadth_first_order, reverse_cuthill_mckee

from oracle import oracle as O


def _meshes():
    from fvmpor_amd.meshgen import read_pmesh
    from fvmpor_amd.problems import box3d_problem, slice2d_problem
    import os
    from conftest import GOLDEN
    return {"cassion": read_pmesh(os.path.join(GOLDEN, "cassion_1_0.pmesh")),
            "slice2d": slice2d_problem(nx=61, ny=25).mesh,
            "box3d": box3d_problem(nx=18, ny=16, nz=8, tiled=False).mesh}


@pytest.fixture(scope="module")
def graphs():
    from fvmpor_amd.learn import node_graph
    return {k: (m, *node_graph(m)) for k, m in _meshes().items()}


def _bandwidth(ia, ja, perm):
    n = len(ia) - 1
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(ia))
    return int(np.abs(inv[rows] - inv[ja]).max())


def _ecc(ia, ja, s):
    n = len(ia) - 1
    A = sp.csr_matrix((np.ones(len(ja)), ja, ia), shape=(n, n))
    from scipy.sparse.csgraph import shortest_path
    return int(shortest_path(A, unweighted=True, indices=[s]).max())


@pytest.mark.parametrize("which", ["cassion", "slice2d", "box3d"])
def test_oracle_sweep_reproduces_scipy(graphs, which):
    m, ia, ja = graphs[which]
    n = len(ia) - 1
    A = sp.csr_matrix((np.ones(len(ja)), ja, ia), shape=(n, n)) + sp.identity(n, format="csr")
    A.sort_indices()
    ref = reverse_cuthill_mckee(A, symmetric_mode=True)
    assert len(breadth_first_order(A, 0, return_predecessors=False)) == n  # connected
    got = O.rcm(ia, ja, seed=int(ref[-1]))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("which", ["cassion", "slice2d", "box3d"])
def test_library_rcm_matches_oracle_and_cuts_bandwidth(graphs, which):
    from fvmpor_amd.lib import rcm_order
    m, ia, ja = graphs[which]
    n = len(ia) - 1
    perm = rcm_order(m)
    orc = O.rcm(ia, ja)
    assert np.array_equal(perm, orc)
    assert np.array_equal(np.sort(perm), np.arange(n))
    start = int(perm[-1])
    deg = np.diff(ia)
    low = int(np.flatnonzero(deg == deg.min())[0])
    assert _ecc(ia, ja, start) >= _ecc(ia, ja, low)
    A = sp.csr_matrix((np.ones(len(ja)), ja, ia), shape=(n, n))
    bw_scipy = _bandwidth(ia, ja, reverse_cuthill_mckee(A, symmetric_mode=True))
    bw = _bandwidth(ia, ja, perm)
    print(f"{which}: bandwidth original {_bandwidth(ia, ja, np.arange(n))}, RCM {bw}, scipy RCM {bw_scipy}")
    shuffled = _bandwidth(ia, ja, np.random.default_rng(1).permutation(n))
    assert bw < shuffled / 4 and bw <= 1.25 * bw_scipy


def test_renumbered_problem_keeps_the_residual(graphs):
    """Renumbering is a relabelling: the oracle residual of the RCM-renumbered
    slice is the permuted residual of the original."""
    from fvmpor_amd.lib import rcm_order
    from fvmpor_amd.meshgen import renumber_nodes
    from fvmpor_amd.problems import slice2d_problem
    p = slice2d_problem(nx=31, ny=13)
    perm = rcm_order(p.mesh)
    q = slice2d_problem(nx=31, ny=13)
    q.mesh = renumber_nodes(p.mesh, perm)
    q.psi0 = p.psi0[perm]
    rng = np.random.default_rng(3)
    psi = p.psi0 + rng.uniform(-0.3, 0.3, len(perm))
    psip = rng.normal(scale=0.1, size=len(perm))
    F = O.OracleProblem(p).residual(0.0, psi, psip)
    G = O.OracleProblem(q).residual(0.0, psi[perm], psip[perm])
    np.testing.assert_allclose(G, F[perm], rtol=1e-12, atol=1e-12 * np.abs(F).max())
