"""What a learned diagonal correction of the block ILU(0) preconditioner can
gain at best, on the CPU with the oracle.

The correction is a per-node diagonal D added to each diagonal block before
its ILU(0), chosen so that the corrected factor is exact on a filter vector t
(D_i = ((A - LU) t)_i / t_i, the tangential / filtering correction).  A network
trained on (state, dt) -> Newton-update pairs can at best supply the true Newton
update as t.  With that best case the corrected preconditioner needs MORE GMRES
iterations than plain block ILU(0) on the Richards Newton systems, at short and
long steps, so the learned path keeps the plain preconditioner."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def _gmres_iterations(A, M, b, m=30, tol=1e-8, maxit=1500):
    """Right-preconditioned restarted GMRES (Givens); iterations to tol ||b||."""
    n = len(b)
    x = np.zeros(n)
    it = 0
    nb = np.linalg.norm(b)
    r = b.copy()
    while it < maxit:
        beta = np.linalg.norm(r)
        Q = np.zeros((n, m + 1))
        H = np.zeros((m + 1, m))
        Q[:, 0] = r / beta
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        for j in range(m):
            w = A @ M(Q[:, j])
            for _ in range(2):
                h = Q[:, :j + 1].T @ w
                H[:j + 1, j] += h
                w = w - Q[:, :j + 1] @ h
            H[j + 1, j] = np.linalg.norm(w)
            Q[:, j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / d, H[j + 1, j] / d
            H[j, j], H[j + 1, j] = d, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            it += 1
            if abs(g[j + 1]) < tol * nb:
                y = np.linalg.solve(np.triu(H[:j + 1, :j + 1]), g[:j + 1])
                return it, x + M(Q[:, :j + 1] @ y)
        y = np.linalg.solve(np.triu(H[:m, :m]), g[:m])
        x = x + M(Q[:, :m] @ y)
        r = b - A @ x
    return it, x


@pytest.fixture(scope="module")
def newton_system():
    from fvmpor_amd.problems import box3d_problem
    prob = box3d_problem(nx=20, ny=20, nz=8)
    op = O.OracleProblem(prob)
    ia, ja = O.jac_pattern(op.omesh)
    return prob, op, np.asarray(ia), np.asarray(ja), O.colouring(ia, ja)


@pytest.mark.parametrize("h", [0.1, 10.0])
def test_filtered_block_ilu_with_the_true_newton_update_does_not_pay(newton_system, h):
    prob, op, ia, ja, (nc, col) = newton_system
    N, bs = op.N, 64
    y, yp = prob.psi0.copy(), np.zeros(N)
    F0 = op.residual(0.0, y, yp)
    ewt = 1.0 / (1e-6 * np.abs(y) + 1e-6)
    Jv = O.fd_jacobian(op, 0.0, 1.0 / h, h, y, yp, F0, ewt, ia, ja, col, nc)
    J = sp.csr_matrix((Jv, ja, ia), shape=(N, N))
    rows = np.repeat(np.arange(N), np.diff(ia))
    keep = rows // bs == ja // bs
    kia = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=N))]).astype(np.int32)
    kja = ja[keep].astype(np.int32)

    def block_ilu(vals):
        lu, dp = O.ilu0(kia, kja, vals[keep])
        F = sp.csr_matrix((lu, kja, kia), shape=(N, N))
        LU = ((sp.tril(F, -1) + sp.eye(N)) @ (sp.triu(F, 1) + sp.diags(1.0 / lu[dp]))).tocsr()
        return (lambda v: O.ilu0_apply(kia, kja, lu, dp, v)), LU

    M0, LU0 = block_ilu(Jv)
    b = -F0
    it0, x = _gmres_iterations(J, M0, b)
    # the filter vector: the true Newton update (floored away from zero)
    floor = 1e-3 * np.abs(x).max()
    t = np.where(np.abs(x) > floor, x, floor)
    D = (J @ t - LU0 @ t) / t
    Jd = Jv.copy()
    Jd[rows == ja] += D
    M1, LU1 = block_ilu(Jd)
    it1, x1 = _gmres_iterations(J, M1, b)
    # the correction does what it is built for: the factor is nearly exact on t
    e0 = np.linalg.norm(J @ t - LU0 @ t) / np.linalg.norm(J @ t)
    e1 = np.linalg.norm(J @ t - LU1 @ t) / np.linalg.norm(J @ t)
    print(f"h={h}: block ILU(0) {it0} iterations, filtered {it1}; |(A - LU) t| / |A t| {e0:.2e} -> {e1:.2e}")
    assert e1 < 0.5 * e0
    assert np.abs(x1 - x).max() < 1e-5 * np.abs(x).max()
    assert it1 >= it0
