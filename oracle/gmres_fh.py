"""TEST INFRASTRUCTURE ONLY: numpy restatement of decomp's deflated restarted
GMRES (gmresFH, gmres.c:212-405; arnoldi, gmres.c:412-511; deflate,
gmres.c:520-849).  The reference does the small dense work with LAPACK (dgesv,
dgeevx through mtx_eigs, QR); numpy calls the same LAPACK routines.  Only
tests/ may import this module.

Right-preconditioned: the Arnoldi basis is built for A M; x += M Q y.  After
each cycle the k harmonic Ritz vectors of smallest magnitude (both real and
imaginary parts of a complex pair) are kept at the front of the next basis.
"""
from __future__ import annotations

import numpy as np


def arnoldi(AM, Q, H, K, m, reorth):
    """gmres.c:412-511: columns K..m-1; returns the number of columns j."""
    j = K
    while j < m:
        w = AM(Q[:, j])
        for i in range(j + 1):
            h = w @ Q[:, i]
            H[i, j] = h
            w = w - h * Q[:, i]
        if reorth:
            for i in range(j + 1):
                h = w @ Q[:, i]
                H[i, j] += h
                w = w - h * Q[:, i]
        h = np.linalg.norm(w)
        H[j + 1, j] = h
        if h < 1e-16:  # lucky breakdown: column j + 1 is not stored
            break
        Q[:, j + 1] = w / h
        j += 1
    return j


def harmonic_ritz(Hbar, k):
    """deflate(), gmres.c:527-651: returns (X with orthonormal columns spanning
    the k smallest-magnitude harmonic Ritz vectors, f, H = Hm + beta^2 f e_m^T,
    beta, eigenvalues)."""
    m = Hbar.shape[1]
    beta = Hbar[m, m - 1]
    Hm = Hbar[:m, :m]
    em = np.zeros(m)
    em[-1] = 1.0
    f = np.linalg.solve(Hm.T, em)
    H = Hm + beta * beta * np.outer(f, em)
    lam, V = np.linalg.eig(H)
    order = np.argsort(np.abs(lam) ** 2, kind="stable")
    cols, i, p = [], 0, 0
    while i < k:
        pos = order[i + p]
        if lam[pos].imag != 0.0:
            v = V[:, pos] if lam[pos].imag > 0 else np.conj(V[:, pos])
            cols += [v.real, v.imag]
            p += 1
        else:
            cols.append(V[:, pos].real)
        i += 1
    X, _ = np.linalg.qr(np.stack(cols, axis=1))
    return X, f, H, beta, lam


def gmres_fh(A, b, x0, M=None, m=20, k=4, max_restarts=50, tol=1e-10, reorth=1):
    """gmresFH(): A, M callables (M None = no preconditioner).  Returns
    (x, dict(restarts, iterations, residuals per restart, converged, K per restart))."""
    n = len(b)
    Mv = (lambda v: v) if M is None else M
    AM = lambda v: A(Mv(v))
    x = np.array(x0, dtype=np.float64)
    Q = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    r = b - A(x)
    beta = np.linalg.norm(r)
    normb = np.linalg.norm(b)
    hist = dict(restarts=0, iterations=0, residuals=[], Ks=[], converged=False)
    if beta == 0.0 or beta / normb < tol:
        hist["converged"] = True
        return x, hist
    Q[:, 0] = r / beta
    K = 0
    restarts = 0
    while restarts <= max_restarts:
        j = arnoldi(AM, Q, H, K, m, reorth)
        hist["iterations"] += j - K
        # least squares min || Q^T r - Hbar y || over the j computed columns
        c = Q[:, : j + 1].T @ r
        y, *_ = np.linalg.lstsq(H[: j + 1, :j], c, rcond=None)
        x = x + Mv(Q[:, :j] @ y)
        r = b - A(x)
        beta = np.linalg.norm(r)
        hist["residuals"].append(beta)
        if beta / normb < tol:
            hist["converged"] = True
            break
        restarts += 1
        if restarts > max_restarts:
            break
        if k and j == m:
            X, f, Hh, hb, _ = harmonic_ritz(H, k)
            jj = X.shape[1]
            Qk = Q[:, :m] @ X
            pm = hb * Q[:, m] - hb * hb * (Q[:, :m] @ f)
            for i in range(jj):  # MGS against Q_k
                pm = pm - (pm @ Qk[:, i]) * Qk[:, i]
            betak = np.linalg.norm(pm)
            pm = pm / betak
            Om = X.T @ Hh @ X - hb * hb * np.outer(X.T @ f, X[m - 1, :])
            H[:] = 0.0
            H[:jj, :jj] = Om
            H[jj, :jj] = betak * X[m - 1, :]
            Q[:] = 0.0
            Q[:, :jj] = Qk
            Q[:, jj] = pm
            K = jj
        else:
            H[:] = 0.0
            Q[:] = 0.0
            Q[:, 0] = r / beta
            K = 0
        hist["Ks"].append(K)
    hist["restarts"] = restarts if hist["converged"] else restarts - 1
    return x, hist
