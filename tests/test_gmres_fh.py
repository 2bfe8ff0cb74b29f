"""Deflated restarted GMRES (gmresFH / deflate, gmres.c:212-405, 527-849).

CPU pins of the numpy restatement (oracle/gmres_fh.py): the decomp known
answer b = A 1 => x = 1; the kept values are harmonic Ritz values, the
eigenvalues of the pencil (Hbar^T Hbar, H_m^T) (scipy); after a deflated
restart the basis satisfies the Arnoldi relation A M Q_K = Q_{K+1} Hbar_K; and
deflation converges where plain restarted GMRES(m) stalls.  The GPU tests run
fvp_gmres_csr_deflated against the restatement on the same systems."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import gmres_fh as G


def _system(n=300, seed=0):
    """Nonsymmetric tridiagonal with five small eigenvalues (they stall GMRES(20))."""
    rng = np.random.default_rng(seed)
    A = sp.diags([-1.3 * np.ones(n - 1), 2.0 * np.ones(n), -0.7 * np.ones(n - 1)], [-1, 0, 1]).tocsr()
    d = np.zeros(n)
    d[:5] = np.array([1e-3, 3e-3, 1e-2, 3e-2, 0.05]) - 1.0
    A = (A + sp.diags(d) + sp.diags(rng.uniform(-0.05, 0.05, n - 2), 2)).tocsr()
    A.sort_indices()
    return A


def test_known_answer_and_gain_over_plain_restarts():
    A = _system()
    n = A.shape[0]
    b = A @ np.ones(n)
    x, h = G.gmres_fh(lambda v: A @ v, b, np.zeros(n), m=20, k=4, max_restarts=300, tol=1e-10)
    assert h["converged"] and np.abs(x - 1.0).max() < 1e-7
    x0, h0 = G.gmres_fh(lambda v: A @ v, b, np.zeros(n), m=20, k=0, max_restarts=300, tol=1e-10)
    print(f"deflated: {h['iterations']} iterations; plain: {h0['iterations']} ({h0['converged']})")
    assert not h0["converged"] or h0["iterations"] > 1.5 * h["iterations"]


def test_harmonic_ritz_values_and_deflated_arnoldi_relation():
    A = _system(200, 1)
    n, m, k = A.shape[0], 20, 4
    AM = lambda v: A @ v
    Q = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    b = np.random.default_rng(2).normal(size=n)
    Q[:, 0] = b / np.linalg.norm(b)
    assert G.arnoldi(AM, Q, H, 0, m, 1) == m
    X, f, Hh, hb, lam = G.harmonic_ritz(H, k)
    pencil = sla.eig(H.T @ H, H[:m, :m].T, right=False)
    np.testing.assert_allclose(np.sort(np.abs(lam)), np.sort(np.abs(pencil)), rtol=1e-8)
    np.testing.assert_allclose(X.T @ X, np.eye(X.shape[1]), atol=1e-12)
    # deflated restart of the reference: Q_K = Q_m X, Hbar_K from Omega; A Q_K = Q_{K+1} Hbar_K
    jj = X.shape[1]
    Qk = Q[:, :m] @ X
    pm = hb * Q[:, m] - hb * hb * (Q[:, :m] @ f)
    for i in range(jj):
        pm = pm - (pm @ Qk[:, i]) * Qk[:, i]
    betak = np.linalg.norm(pm)
    Qn = np.column_stack([Qk, pm / betak])
    Hk = np.zeros((jj + 1, jj))
    Hk[:jj] = X.T @ Hh @ X - hb * hb * np.outer(X.T @ f, X[m - 1, :])
    Hk[jj] = betak * X[m - 1, :]
    lhs = np.column_stack([A @ Qk[:, i] for i in range(jj)])
    assert np.abs(lhs - Qn @ Hk).max() < 1e-10 * np.abs(lhs).max()


def _csr_dev(torch, A):
    from fvmpor_amd.lib import CsrMatrix
    M = CsrMatrix(A.indptr.astype(np.int32), A.indices.astype(np.int32))
    M.set_values(torch.tensor(A.data, dtype=torch.float64, device="cuda:0"))
    return M


@pytest.mark.gpu
@pytest.mark.parametrize("k", [0, 4])
@pytest.mark.parametrize("jacobi", [False, True])
def test_device_matches_restatement(k, jacobi):
    """Eight restarts of both: the same Arnoldi step count and the same true
    residual (rounding apart); then a run to convergence reaches x = 1 with
    an iteration count near the restatement's (the restart paths part
    through rounding over tens of deflations)."""
    import torch
    from conftest import device
    from fvmpor_amd.lib import PRECON_JACOBI, PRECON_NONE
    device()
    A = _system()
    n = A.shape[0]
    b = A @ np.ones(n)
    Mdev = _csr_dev(torch, A)
    assert Mdev.precon_build(PRECON_JACOBI if jacobi else PRECON_NONE) == 0
    Dinv = 1.0 / A.diagonal()
    M = (lambda v: Dinv * v) if jacobi else None
    tb = torch.tensor(b, device="cuda:0")
    x = torch.zeros_like(tb)
    rc, st = Mdev.gmres_deflated(tb, x, k=k, m=20, max_restarts=8, tol=1e-14)
    xo, ho = G.gmres_fh(lambda v: A @ v, b, np.zeros(n), M=M, m=20, k=k, max_restarts=8, tol=1e-14)
    torch.cuda.synchronize()
    rel = abs(st["residual"] - ho["residuals"][-1]) / ho["residuals"][-1]
    print(f"k={k} jacobi={jacobi}: device {st}, oracle iterations {ho['iterations']}, residual rel diff {rel:.2e}")
    assert st["iterations"] == ho["iterations"]
    assert rel < 1e-5
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-6, atol=1e-6 * np.abs(xo).max())
    x.zero_()
    rc, st = Mdev.gmres_deflated(tb, x, k=k, m=20, max_restarts=400, tol=1e-10)
    xo, ho = G.gmres_fh(lambda v: A @ v, b, np.zeros(n), M=M, m=20, k=k, max_restarts=400, tol=1e-10)
    torch.cuda.synchronize()
    print(f"   to convergence: device {st['iterations']} iterations ({st['converged']}), "
          f"oracle {ho['iterations']} ({ho['converged']})")
    assert bool(st["converged"]) == ho["converged"]
    if ho["converged"]:
        assert np.abs(x.cpu().numpy() - 1.0).max() < 1e-6
        assert abs(st["iterations"] - ho["iterations"]) <= 0.2 * ho["iterations"]
