"""TEST INFRASTRUCTURE ONLY: numpy restatement of decomp's one-level Schur
complement preconditioner on a single process.  Only tests/ may import it.

decomp (schur.c) splits each subdomain's unknowns into interior ones, all of
whose couplings stay inside the subdomain, and boundary (interface) ones
(mtx_CRS_dist_split, schur.c:82-600), orders them interior first, and forms

    A = [ B  E ]   B: interior x interior (block diagonal over subdomains)
        [ F  C ]   C: boundary x boundary (couples subdomains)

precon_schur (schur.c:1247-1323) builds the B preconditioner M_B (here ILU(0)
of each B_i) and the local Schur complements column by column,
S_i = C_ii - F_i M_B^-1 E_i (schur_form_local, schur.c:887-981: a dense column
of E, precon_apply, -F times it, plus the C column).  The level-0 preconditioner
on S is block-Jacobi with the exact inverse of each dense S_i: the couplings of
S between subdomains (C_ij, i != j) are dropped.  precon_schur_apply
(schur.c:1583-1626) is

    u  = M_B^-1 x_I
    yG = S^-1 (x_G - F u)
    yI = M_B^-1 (x_I - E yG)

Subdomains are the consecutive blocks of `bs` rows; a row is a boundary row
when it has an entry in another block or another block has an entry in its
column.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import oracle as O


def split(ia, ja, bs):
    """(interior rows, boundary rows) per subdomain, each ascending."""
    n = len(ia) - 1
    sub = np.arange(n) // bs
    rows = np.repeat(np.arange(n), np.diff(ia))
    cross = sub[rows] != sub[np.asarray(ja)]
    bnd = np.zeros(n, dtype=bool)
    bnd[rows[cross]] = True
    bnd[np.asarray(ja)[cross]] = True
    nsub = (n + bs - 1) // bs
    I = [np.flatnonzero(~bnd[b * bs:(b + 1) * bs]) + b * bs for b in range(nsub)]
    G = [np.flatnonzero(bnd[b * bs:(b + 1) * bs]) + b * bs for b in range(nsub)]
    return I, G


def gauss_jordan_inverse(S):
    """In-place Gauss-Jordan inversion with partial (row) pivoting, the
    interchanges undone as column swaps in reverse order at the end.  The
    first maximum of |pivot column| is the pivot."""
    a = np.array(S, dtype=np.float64)
    n = a.shape[0]
    perm = np.arange(n)
    for k in range(n):
        p = k + int(np.argmax(np.abs(a[k:, k])))
        if a[p, k] == 0.0:
            raise ZeroDivisionError("singular Schur block")
        perm[k] = p
        if p != k:
            a[[k, p]] = a[[p, k]]
        inv = 1.0 / a[k, k]
        rk = a[k] * inv
        rk[k] = inv
        f = a[:, k].copy()
        a -= np.outer(f, rk)
        a[:, k] = -f * inv
        a[k] = rk
    for k in range(n - 1, -1, -1):
        if perm[k] != k:
            a[:, [k, perm[k]]] = a[:, [perm[k], k]]
    return a


class Schur:
    """The preconditioner's parts for CSR (ia, ja, a) and subdomains of bs rows."""

    def __init__(self, ia, ja, a, bs):
        n = len(ia) - 1
        A = sp.csr_matrix((np.asarray(a, dtype=np.float64), np.asarray(ja), np.asarray(ia)), shape=(n, n))
        self.I, self.G = split(ia, ja, bs)
        self.Iall = np.concatenate(self.I).astype(np.int64)
        self.Gall = np.concatenate(self.G).astype(np.int64)
        self.goff = np.cumsum([0] + [len(g) for g in self.G])
        B = A[self.Iall][:, self.Iall].tocsr()
        B.sort_indices()
        self.bia, self.bja = B.indptr.astype(np.int32), B.indices.astype(np.int32)
        self.blu, self.bdp = O.ilu0(self.bia, self.bja, B.data)
        if self.blu is None:
            raise ZeroDivisionError("zero pivot in the interior ILU(0)")
        self.E = A[self.Iall][:, self.Gall].tocsr()
        self.F = A[self.Gall][:, self.Iall].tocsr()
        C = A[self.Gall][:, self.Gall].tocsr()
        # local Schur complements, one column index at a time for every
        # subdomain at once (M_B is block diagonal, so the columns of different
        # subdomains do not mix)
        ng = [len(g) for g in self.G]
        maxg = max(ng) if ng else 0
        self.S = [np.zeros((m, m)) for m in ng]
        Ed = self.E.toarray()
        for c in range(maxg):
            t = np.zeros(len(self.Iall))
            for b, m in enumerate(ng):
                if c < m:
                    t += Ed[:, self.goff[b] + c]
            y = self.mb(t)
            fy = self.F @ y
            for b, m in enumerate(ng):
                if c < m:
                    self.S[b][:, c] = -fy[self.goff[b]:self.goff[b + 1]]
        for b, m in enumerate(ng):
            lo, hi = self.goff[b], self.goff[b + 1]
            self.S[b] += C[lo:hi, lo:hi].toarray()
        self.Sinv = [gauss_jordan_inverse(s) for s in self.S]

    def mb(self, v):
        return O.ilu0_apply(self.bia, self.bja, self.blu, self.bdp, v)

    def apply(self, x):
        x = np.asarray(x, dtype=np.float64)
        xI, xG = x[self.Iall], x[self.Gall]
        u = self.mb(xI)
        g = xG - self.F @ u
        yG = np.concatenate([self.Sinv[b] @ g[self.goff[b]:self.goff[b + 1]] for b in range(len(self.G))])
        yI = self.mb(xI - self.E @ yG)
        y = np.empty_like(x)
        y[self.Iall] = yI
        y[self.Gall] = yG
        return y
