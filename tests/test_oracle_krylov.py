"""Pins of the oracle's linear-solver restatement (oracle/orc_krylov.c): the
decomp known answer (b = A 1 => x = 1, solver.c:151-166), exact solves for
m >= n against numpy, Arnoldi relations, ILU(0) identities, the coloured FD
Jacobian against J v."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def _csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def _cassion_matrix(cassion_mesh, seed=0, shift=4.0):
    om = O.OracleMesh(cassion_mesh)
    ia, ja = O.jac_pattern(om)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, len(ja))
    n = len(ia) - 1
    for i in range(n):
        for k in range(ia[i], ia[i + 1]):
            if ja[k] == i:
                a[k] = shift * (ia[i + 1] - ia[i])
    return ia, ja, a


def test_csr_gemv_matches_scipy(cassion_mesh):
    ia, ja, a = _cassion_matrix(cassion_mesh)
    x = np.random.default_rng(1).normal(size=len(ia) - 1)
    A = sp.csr_matrix((a, ja, ia))
    np.testing.assert_allclose(O.csr_gemv(ia, ja, a, x), A @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("precon", [0, 2, 5])
def test_gmres_known_answer(cassion_mesh, precon):
    ia, ja, a = _cassion_matrix(cassion_mesh, shift=1.5)
    n = len(ia) - 1
    b = O.csr_gemv(ia, ja, a, np.ones(n))
    # m = 10: gmres.c projects the residual explicitly (y = beta Q^T r, gmres.c:101), which
    # loses accuracy once a cycle runs far past convergence; restarts keep the cycles short
    x, st = O.gmres_csr(ia, ja, a, b, precon=precon, m=10, max_restarts=50, tol=1e-12)
    assert st["converged"]
    assert np.abs(x - 1).max() < 1e-9


def test_gmres_full_krylov_is_exact():
    rng = np.random.default_rng(2)
    n = 60
    A = rng.normal(size=(n, n)) + 3 * np.eye(n)
    ia, ja, a = _csr(A)
    b = rng.normal(size=n)
    # a full Krylov space ends in a lucky breakdown whose last column gmres.c drops
    # (arnoldi returns j without it, gmres.c:485-491), so each cycle is GMRES(n-1)
    x, st = O.gmres_csr(ia, ja, a, b, m=n, max_restarts=10, tol=1e-13, reorth=1)
    assert st["converged"]
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)



@pytest.mark.parametrize("precon", [0, 2, 5])
def test_gmres_does_not_depend_on_heap_contents(precon):
    """The CRS gemv scales its output by beta = 0 before summing, as the
    reference's does, so a work vector taken uninitialised from the heap could
    carry a NaN into GMRES.  Freed heap chunks of the work vectors' size are
    filled with NaN bytes (0xFF) right before each solve; the solve must still
    converge to the known answer."""
    import ctypes
    n = 1000  # 8 KB vectors: below the allocator's mmap threshold, so freed chunks are reused
    rng = np.random.default_rng(1)
    rows = [np.unique(np.concatenate([[i], rng.integers(0, n, 6)])) for i in range(n)]
    ia = np.zeros(n + 1, dtype=np.int32)
    ia[1:] = np.cumsum([len(r) for r in rows])
    ja = np.concatenate(rows).astype(np.int32)
    a = rng.uniform(-1, 1, len(ja))
    a[ja == np.repeat(np.arange(n), np.diff(ia))] = 10.0
    b = O.csr_gemv(ia, ja, a, np.ones(n))
    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]
    libc.free.argtypes = [ctypes.c_void_p]
    for _ in range(3):
        chunks = [libc.malloc(8 * n) for _ in range(32)]
        for c in chunks:
            ctypes.memset(c, 0xFF, 8 * n)
        for c in chunks:
            libc.free(c)
        x, st = O.gmres_csr(ia, ja, a, b, precon=precon, m=10, max_restarts=50, tol=1e-12)
        assert st["converged"], st
        assert np.abs(x - 1.0).max() < 1e-9

@pytest.mark.parametrize("reorth", [0, 1])
def test_arnoldi_relations(cassion_mesh, reorth):
    ia, ja, a = _cassion_matrix(cassion_mesh, shift=1.0)
    n = len(ia) - 1
    A = sp.csr_matrix((a, ja, ia))
    b = np.random.default_rng(3).normal(size=n)
    m = 25 if reorth else 8  # MGS alone loses orthogonality as the Krylov residual falls
    j, Q, H = O.arnoldi_csr(ia, ja, a, b, m, reorth)
    assert j == m
    np.testing.assert_allclose(Q.T @ Q, np.eye(m + 1), atol=1e-10 if reorth else 1e-8)
    np.testing.assert_allclose(A @ Q[:, :m], Q @ H, atol=1e-10)
    # H is upper Hessenberg
    assert np.abs(np.tril(H, -2)).max() == 0


def test_ilu0_is_exact_lu_on_tridiagonal():
    n = 50
    rng = np.random.default_rng(4)
    A = sp.diags([rng.uniform(-1, 1, n - 1), 4 + rng.uniform(0, 1, n), rng.uniform(-1, 1, n - 1)], [-1, 0, 1])
    ia, ja, a = _csr(A)
    lu, dp = O.ilu0(ia, ja, a)
    x = rng.normal(size=n)
    np.testing.assert_allclose(O.ilu0_apply(ia, ja, lu, dp, x), np.linalg.solve(A.toarray(), x), rtol=1e-12)


def test_ilu0_reproduces_pattern_entries(cassion_mesh):
    ia, ja, a = _cassion_matrix(cassion_mesh)
    n = len(ia) - 1
    lu, dp = O.ilu0(ia, ja, a)
    L = sp.lil_matrix((n, n))
    U = sp.lil_matrix((n, n))
    for i in range(n):
        L[i, i] = 1.0
        for k in range(ia[i], ia[i + 1]):
            j = ja[k]
            if j < i:
                L[i, j] = lu[k]
            elif j == i:
                U[i, i] = 1.0 / lu[k]
            else:
                U[i, j] = lu[k]
    P = (L.tocsr() @ U.tocsr()).tocsr()
    A = sp.csr_matrix((a, ja, ia))
    rows = np.repeat(np.arange(n), np.diff(ia))
    np.testing.assert_allclose(np.asarray(P[rows, ja]).ravel(), np.asarray(A[rows, ja]).ravel(), atol=1e-11)


def test_jacobi_and_zero_pivot():
    A = sp.csr_matrix(np.array([[2.0, 1.0], [0.0, 4.0]]))
    ia, ja, a = _csr(A)
    np.testing.assert_allclose(O.jacobi(ia, ja, a), [0.5, 0.25])
    B = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 0.0]]))
    B.eliminate_zeros()
    ib, jb, bb = _csr(B)
    assert O.jacobi(ib, jb, bb) is None
    assert O.ilu0(ib, jb, bb)[0] is None


def test_colouring_is_distance_two(cassion_mesh):
    om = O.OracleMesh(cassion_mesh)
    ia, ja = O.jac_pattern(om)
    nc, col = O.colouring(ia, ja)
    n = len(ia) - 1
    for i in range(n):
        seen = {}
        for j in ja[ia[i]:ia[i + 1]]:
            c = col[j]
            assert c not in seen, (i, j, seen[c])
            seen[c] = j
    assert nc == col.max() + 1


@pytest.mark.parametrize("which", ["slice2d", "box3d_small"])
def test_fd_jacobian_matches_jv(which, request):
    prob = request.getfixturevalue(which)
    op = O.OracleProblem(prob)
    rng = np.random.default_rng(7)
    y = prob.psi0 - 0.3 + rng.uniform(-0.2, 0.2, op.N)
    yp = rng.normal(scale=0.05, size=op.N)
    op.begin_step(0.0, y, yp)
    F0 = op.residual(0.0, y, yp)
    ia, ja = O.jac_pattern(op.omesh)
    nc, col = O.colouring(ia, ja)
    ewt = 1.0 / (1e-6 * np.abs(y) + 1e-6)
    c, h = 12.0, 0.05
    J = O.fd_jacobian(op, 0.0, c, h, y, yp, F0, ewt, ia, ja, col, nc)
    v = rng.normal(size=op.N)
    Jv = O.jv(op, 0.0, y, yp, F0, c, v)
    Jm = sp.csr_matrix((J, ja, ia))
    err = np.abs(-(Jm @ v) - Jv).max() / np.abs(Jv).max()
    assert err < 1e-5, err
