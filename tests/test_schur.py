"""The oracle's Schur-complement preconditioner (schur.c's one-level form on
one process) against what the mathematics fixes: Gauss-Jordan against LAPACK's
inverse, the exact inverse when M_B is exact, block ILU(0) when no row is on a
subdomain boundary, and S_i against the dense product C_ii - F_i M_B^-1 E_i."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from oracle.schur_ref import Schur, gauss_jordan_inverse, split


def _mesh_matrix(seed=0, nx=9, ny=8, nz=5):
    from fvmpor_amd.problems import box3d_problem
    prob = box3d_problem(nx=nx, ny=ny, nz=nz)
    ia, ja = O.jac_pattern(O.OracleMesh(prob.mesh))
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, len(ja))
    rows = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
    a[rows == ja] = 1.3 * np.diff(ia)[rows[rows == ja]]
    return np.asarray(ia, np.int32), np.asarray(ja, np.int32), a


def test_gauss_jordan_matches_lapack():
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 64, 128):
        S = rng.normal(size=(n, n))
        ref = np.linalg.inv(S)
        assert np.abs(gauss_jordan_inverse(S) - ref).max() <= 1e-11 * np.abs(ref).max()
    # needs every interchange: a permutation's inverse is its transpose
    P = np.eye(7)[[3, 6, 1, 5, 0, 2, 4]]
    assert np.array_equal(gauss_jordan_inverse(P), P.T)
    with pytest.raises(ZeroDivisionError):
        gauss_jordan_inverse(np.ones((3, 3)))


def test_split_marks_both_ends_of_cross_couplings():
    # 3 x 3 blocks of a chain with one extra coupling 0 -> 7 stored in row 0 only
    n, bs = 9, 3
    A = sp.diags([np.ones(n - 1), 4 * np.ones(n), np.ones(n - 1)], [-1, 0, 1], format="lil")
    A[0, 7] = 0.5
    A = A.tocsr()
    I, G = split(A.indptr, A.indices, bs)
    assert [list(g) for g in G] == [[0, 2], [3, 5], [6, 7]]
    assert [list(i) for i in I] == [[1], [4], [8]]


def test_exact_when_interior_blocks_are_diagonal():
    """With B diagonal, ILU(0) of B is B itself, and S_i = C_ii - F_i B^-1 E_i is
    the exact Schur complement of A without the couplings between different
    subdomains' boundary rows; the preconditioner then inverts that matrix."""
    ia, ja, a = _mesh_matrix()
    n, bs = len(ia) - 1, 64
    I, G = split(ia, ja, bs)
    inter = np.zeros(n, bool)
    inter[np.concatenate(I)] = True
    rows = np.repeat(np.arange(n), np.diff(ia))
    a = a.copy()
    a[inter[rows] & inter[ja] & (rows != ja)] = 0.0          # B diagonal
    a[~inter[rows] & ~inter[ja] & (rows // bs != ja // bs)] = 0.0  # no C_ij, i != j
    A = sp.csr_matrix((a, ja, ia), shape=(n, n))
    assert sum(len(g) for g in G) > n // 3 and max(len(g) for g in G) > 10
    P = Schur(ia, ja, a, bs)
    x = np.random.default_rng(2).normal(size=n)
    np.testing.assert_allclose(P.apply(A @ x), x, rtol=0, atol=1e-11 * np.abs(x).max())


def test_no_boundary_rows_is_block_ilu0():
    ia, ja, a = _mesh_matrix(3)
    n, bs = len(ia) - 1, 128
    rows = np.repeat(np.arange(n), np.diff(ia))
    keep = rows // bs == ja // bs
    kia = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int32)
    kja, ka = ja[keep], a[keep]
    P = Schur(kia, kja, ka, bs)
    assert len(P.Gall) == 0
    lu, dp = O.ilu0(kia, kja, ka)
    x = np.random.default_rng(4).normal(size=n)
    np.testing.assert_array_equal(P.apply(x), O.ilu0_apply(kia, kja, lu, dp, x))


def test_local_schur_blocks_are_the_dense_products():
    """Each S_i against C_ii - F_i M_B^-1 E_i with M_B^-1 as a dense matrix (its
    columns are M_B^-1 applied to unit vectors): checks that forming one column
    index of every subdomain at once does not mix subdomains."""
    ia, ja, a = _mesh_matrix(5)
    n, bs = len(ia) - 1, 64
    P = Schur(ia, ja, a, bs)
    nI = len(P.Iall)
    Minv = np.stack([P.mb(e) for e in np.eye(nI)], axis=1)
    A = sp.csr_matrix((a, ja, ia), shape=(n, n)).toarray()
    for b in range(len(P.G)):
        g, lo, hi = P.G[b], P.goff[b], P.goff[b + 1]
        Fb = A[np.ix_(g, P.Iall)]
        Eb = A[np.ix_(P.Iall, g)]
        ref = A[np.ix_(g, g)] - Fb @ Minv @ Eb
        assert hi - lo == len(g)
        np.testing.assert_allclose(P.S[b], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        np.testing.assert_allclose(P.Sinv[b] @ P.S[b], np.eye(len(g)), atol=1e-10)
