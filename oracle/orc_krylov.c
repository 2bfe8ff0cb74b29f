/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle: restatement of the reference's
 * linear-solver arithmetic (decomp/) and of the coloured finite-difference
 * Jacobian of its direct-solve preconditioner.  Only the tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it.
 *
 * Followed reference files (paths relative to the reference root):
 *   decomp/gmres.c:26-201, 412-511   restarted right-preconditioned GMRES, MGS Arnoldi
 *   decomp/linalg_dense.c:128-176    Givens reduction of the Hessenberg matrix
 *   decomp/linalg_sparse.c:1321-1348 CRS gemv (non-transposed, scalar)
 *   decomp/linalg_mpi.c:1016-1031    residual r = b - A x
 *   decomp/ILU.c:12-115, 133-223     ILU(0) factor and apply.  The factor follows
 *       the block path's semantics (diagonal of row i inverted after row i,
 *       ILU.c:108-110); the scalar path's stale index at ILU.c:62 is a
 *       reference defect and is not reproduced.
 *   decomp/MR.c:143-230              Jacobi (inverse diagonal)
 *   examples/fvmporMinlin/preconditioner_dss.cpp:21-94   distance-2 greedy colouring
 *   examples/fvmporMinlin/preconditioner_dss.cpp:107-142 element-pair pattern
 *   examples/fvmporMinlin/preconditioner_dss.cpp:284-372 shifted-residual Jacobian
 *   include/fvm/integrators/eem_integrator.h:323-337      FD J v increment
 * BLAS drotg/drot follow the reference BLAS definitions (netlib).
 *
 * Parity pins: tests/test_oracle_krylov.py (A x = A 1 known answer, m >= n
 * exact solve against numpy, Arnoldi orthonormality, ILU(0) = LU on tree
 * patterns, (LU)_ij = A_ij on the pattern, FD Jacobian against Jv).
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- CRS ------------------------------------------------------------- */
void orc_csr_gemv(int n, const int *ia, const int *ja, const double *a, const double *x, double *y) {
    for (int row = 0; row < n; ++row) {
        y[row] *= 0.;
        for (int k = ia[row]; k < ia[row + 1]; ++k) y[row] += 1. * a[k] * x[ja[k]];
    }
}

/* ILU(0) in place; returns 1 on success, 0 on a missing diagonal (ILU.c:21-30) */
int orc_ilu0_factor(int n, const int *ia, const int *ja, double *nz, int *diagpos) {
    for (int i = 0; i < n; ++i) {
        int k = ia[i];
        while (k < ia[i + 1] && ja[k] != i) k++;
        if (k == ia[i + 1]) return 0;
        diagpos[i] = k;
    }
    for (int i = 0; i < n; ++i) {
        for (int kk = ia[i]; kk < diagpos[i]; ++kk) {
            int k = ja[kk];
            int kpos = ia[k];
            nz[kk] *= nz[diagpos[k]];
            for (int jj = kk + 1; jj < ia[i + 1]; ++jj) {
                int j = ja[jj];
                while (kpos < ia[k + 1] && ja[kpos] < j) kpos++;
                if (kpos == ia[k + 1]) break;
                if (ja[kpos] == j) nz[jj] -= nz[kk] * nz[kpos];
            }
        }
        nz[diagpos[i]] = 1. / nz[diagpos[i]];
    }
    return 1;
}

void orc_ilu0_apply(int n, const int *ia, const int *ja, const double *nz, const int *diagpos, const double *x,
                    double *y) {
    memcpy(y, x, sizeof(double) * n);
    for (int i = 1; i < n; ++i)
        for (int j = ia[i]; j < diagpos[i]; ++j) y[i] -= nz[j] * y[ja[j]];
    for (int i = n - 1; i >= 0; --i) {
        for (int j = diagpos[i] + 1; j < ia[i + 1]; ++j) y[i] -= nz[j] * y[ja[j]];
        y[i] *= nz[diagpos[i]];
    }
}

/* Jacobi: J = 1 / diag(A); 0 on a zero or missing diagonal (MR.c:150-171) */
int orc_jacobi_factor(int n, const int *ia, const int *ja, const double *nz, double *J) {
    for (int i = 0; i < n; ++i) {
        int pos = ia[i];
        while (pos < ia[i + 1] && ja[pos] < i) pos++;
        if (pos == ia[i + 1] || !nz[pos]) return 0;
        J[i] = 1. / nz[pos];
    }
    return 1;
}

/* ---- BLAS helpers ------------------------------------------------------ */
static double dnrm2(int n, const double *x) {
    double s = 0.;
    for (int i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s);
}
static double ddot(int n, const double *x, const double *y) {
    double s = 0.;
    for (int i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}
/* netlib drotg */
static void drotg(double *da, double *db, double *c, double *s) {
    double roe = fabs(*da) > fabs(*db) ? *da : *db;
    double scale = fabs(*da) + fabs(*db);
    double r, z;
    if (scale == 0.) { *c = 1.; *s = 0.; r = 0.; z = 0.; }
    else {
        double a = *da / scale, b = *db / scale;
        r = scale * sqrt(a * a + b * b);
        r = (roe >= 0 ? 1. : -1.) * r;
        *c = *da / r;
        *s = *db / r;
        z = 1.;
        if (fabs(*da) > fabs(*db)) z = *s;
        if (fabs(*db) >= fabs(*da) && *c != 0.) z = 1. / *c;
    }
    *da = r;
    *db = z;
}
static void drot(int n, double *x, int incx, double *y, int incy, double c, double s) {
    for (int i = 0; i < n; ++i) {
        double t = c * x[i * incx] + s * y[i * incy];
        y[i * incy] = c * y[i * incy] - s * x[i * incx];
        x[i * incx] = t;
    }
}

/* ---- GMRES ------------------------------------------------------------- */
typedef void (*OpFn)(void *ctx, const double *x, double *y);

typedef struct {
    int dim_krylov, max_restarts, reorthogonalise;
    double tol;
    /* outputs */
    int restarts, total_iters, converged;
    double residual, error;
} OrcGmres;

/* gmres.c:26-201 with arnoldi (gmres.c:412-511); A and M are operators; M may be NULL */
void orc_gmres(int n, OpFn A, void *actx, OpFn M, void *mctx, const double *b, double *x, OrcGmres *run) {
    int m = run->dim_krylov;
    int lda = m + 1;
    double *Q = (double *)calloc((size_t)n * (m + 1), sizeof(double));
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double)); /* column-major (m+1) x m */
    double *Ht = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *y = (double *)calloc(m + 1, sizeof(double));
    double *r = (double *)calloc(n, sizeof(double));
    double *w = (double *)calloc(n, sizeof(double));
    double *dx = (double *)calloc(n, sizeof(double));
    double *wtmp = (double *)calloc(n, sizeof(double));
    int restarts = 0, converged = 0, total = 0;

    /* r = b - A x */
    A(actx, x, r);
    for (int i = 0; i < n; ++i) r[i] = -r[i] + b[i];
    double beta = dnrm2(n, r), normb = dnrm2(n, b), error;
    run->restarts = 0;
    if (!(error = beta) || (error = beta / normb) < run->tol) {
        run->converged = 1; run->residual = beta; run->error = error; run->total_iters = 0;
        goto done;
    }
    for (int i = 0; i < n; ++i) r[i] *= 1. / beta;
    memcpy(Q, r, sizeof(double) * n);
    while (restarts <= run->max_restarts && !converged) {
        /* Arnoldi */
        int j = 0;
        while (j < m) {
            const double *qj = Q + (size_t)n * j;
            if (M) { M(mctx, qj, wtmp); A(actx, wtmp, w); }
            else A(actx, qj, w);
            int pos = j * lda;
            for (int i = 0; i <= j; ++i, ++pos) {
                const double *qi = Q + (size_t)n * i;
                double h = ddot(n, w, qi);
                H[pos] = h;
                for (int k = 0; k < n; ++k) w[k] += -h * qi[k];
            }
            if (run->reorthogonalise) {
                pos = j * lda;
                for (int i = 0; i <= j; ++i, ++pos) {
                    const double *qi = Q + (size_t)n * i;
                    double h = ddot(n, w, qi);
                    H[pos] += h;
                    for (int k = 0; k < n; ++k) w[k] += -h * qi[k];
                }
            }
            double h = dnrm2(n, w);
            H[pos] = h;
            if (h < 1e-16) break;
            double *qn = Q + (size_t)n * (j + 1);
            for (int k = 0; k < n; ++k) qn[k] += 1. / h * w[k];
            ++j;
        }
        total += j;
        /* y = beta * Q(:,0:j)' r */
        for (int i = 0; i <= j && i <= m; ++i) y[i] = beta * ddot(n, Q + (size_t)n * i, r);
        for (int i = j + 1; i <= m; ++i) y[i] = 0.;
        memcpy(Ht, H, sizeof(double) * (m + 1) * m);
        /* Givens (linalg_dense.c:161-169) */
        for (int i = 0, pos = 0; i < m; ++i, pos += lda) {
            double a = Ht[pos + i], bb = Ht[pos + i + 1], c, s;
            drotg(&a, &bb, &c, &s);
            drot(m - i, &Ht[pos + i], lda, &Ht[pos + i + 1], lda, c, s);
            drot(1, y + i, lda, y + i + 1, lda, c, s);
        }
        /* upper-triangular j x j solve (dtrsm L,U,N,N) */
        for (int i = j - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < j; ++k) s -= Ht[k * lda + i] * y[k];
            y[i] = s / Ht[i * lda + i];
        }
        /* w = Q(:,0:j-1) y ; dx = M w ; x += dx */
        for (int k = 0; k < n; ++k) w[k] = 0.;
        for (int i = 0; i < j; ++i) {
            const double *qi = Q + (size_t)n * i;
            for (int k = 0; k < n; ++k) w[k] += qi[k] * y[i];
        }
        if (M) M(mctx, w, dx);
        else memcpy(dx, w, sizeof(double) * n);
        for (int k = 0; k < n; ++k) x[k] += dx[k];
        A(actx, x, r);
        for (int i = 0; i < n; ++i) r[i] = -r[i] + b[i];
        beta = dnrm2(n, r);
        error = beta / normb;
        if (error < run->tol) converged = 1;
        else restarts++;
        if (!converged && restarts <= run->max_restarts) {
            memset(H, 0, sizeof(double) * (m + 1) * m);
            memset(Q, 0, sizeof(double) * (size_t)n * (m + 1));
            for (int i = 0; i < n; ++i) r[i] *= 1. / beta;
            memcpy(Q, r, sizeof(double) * n);
        }
    }
    run->restarts = converged ? restarts : restarts - 1;
    run->converged = converged;
    run->residual = beta;
    run->error = error;
    run->total_iters = total;
done:
    free(Q); free(H); free(Ht); free(y); free(r); free(w); free(dx); free(wtmp);
}

/* ---- explicit CSR systems (the decomp solver harness, solver.c:151-166) ---- */
typedef struct { int n; const int *ia, *ja; const double *a; } CsrOp;
static void csr_op(void *c, const double *x, double *y) {
    CsrOp *o = (CsrOp *)c;
    orc_csr_gemv(o->n, o->ia, o->ja, o->a, x, y);
}
typedef struct { int n, type; const int *ia, *ja; double *lu; int *diagpos; double *J; } PcOp;
static void pc_op(void *c, const double *x, double *y) {
    PcOp *p = (PcOp *)c;
    if (p->type == 5) orc_ilu0_apply(p->n, p->ia, p->ja, p->lu, p->diagpos, x, y);
    else if (p->type == 2) for (int i = 0; i < p->n; ++i) y[i] = x[i] * p->J[i];
    else memcpy(y, x, sizeof(double) * p->n);
}

/* precon types as precon.h:8-13: 0 none, 2 Jacobi, 5 ILU(0).  Returns 1 on success.
 * iparams: dim_krylov, max_restarts, reorthogonalise; outputs: stats[0..4] =
 * restarts, total iterations, converged, residual, relative error */
int orc_gmres_csr(int n, const int *ia, const int *ja, const double *a, int precon_type, const double *b, double *x,
                  const int *iparams, double tol, double *stats) {
    CsrOp A = {n, ia, ja, a};
    PcOp P = {n, precon_type, ia, ja, NULL, NULL, NULL};
    int ok = 1;
    if (precon_type == 5) {
        P.lu = (double *)calloc(ia[n], sizeof(double));
        memcpy(P.lu, a, sizeof(double) * ia[n]);
        P.diagpos = (int *)malloc(sizeof(int) * n);
        ok = orc_ilu0_factor(n, ia, ja, P.lu, P.diagpos);
    } else if (precon_type == 2) {
        P.J = (double *)calloc(n, sizeof(double));
        ok = orc_jacobi_factor(n, ia, ja, a, P.J);
    }
    if (ok) {
        OrcGmres run = {iparams[0], iparams[1], iparams[2], tol, 0, 0, 0, 0., 0.};
        orc_gmres(n, csr_op, &A, precon_type ? pc_op : NULL, &P, b, x, &run);
        stats[0] = run.restarts; stats[1] = run.total_iters; stats[2] = run.converged;
        stats[3] = run.residual; stats[4] = run.error;
    }
    free(P.lu); free(P.diagpos); free(P.J);
    return ok;
}

/* Arnoldi basis check helper: run j steps of unpreconditioned Arnoldi on CSR A
 * from b, return Q (n x (j+1)) and H ((j+1) x j column-major). */
int orc_arnoldi_csr(int n, const int *ia, const int *ja, const double *a, const double *b, int m, int reorth,
                    double *Q, double *H) {
    CsrOp A = {n, ia, ja, a};
    double beta = dnrm2(n, b);
    memset(Q, 0, sizeof(double) * (size_t)n * (m + 1));
    memset(H, 0, sizeof(double) * (size_t)(m + 1) * m);
    for (int i = 0; i < n; ++i) Q[i] = b[i] / beta;
    double *w = (double *)calloc(n, sizeof(double));
    int j = 0;
    while (j < m) {
        csr_op(&A, Q + (size_t)n * j, w);
        int pos = j * (m + 1);
        for (int pass = 0; pass < (reorth ? 2 : 1); ++pass) {
            pos = j * (m + 1);
            for (int i = 0; i <= j; ++i, ++pos) {
                double h = ddot(n, w, Q + (size_t)n * i);
                H[pos] += h;
                for (int k = 0; k < n; ++k) w[k] -= h * Q[(size_t)n * i + k];
            }
        }
        double h = dnrm2(n, w);
        H[pos] = h;
        if (h < 1e-16) break;
        for (int k = 0; k < n; ++k) Q[(size_t)n * (j + 1) + k] = w[k] / h;
        ++j;
    }
    free(w);
    return j;
}

/* ---- coloured FD Jacobian of the problem (preconditioner_dss.cpp) ------- */
typedef struct OrcProblem OrcProblem;
void orc_residual(OrcProblem *p, double t, const double *h, const double *hp, double *res);

/* node adjacency incl. diagonal from element node pairs; ia (N+1), ja (ia[N]) */
int orc_jac_pattern(const OrcMesh *m, int *ia, int *ja) {
    int N = m->N, npe = m->npe;
    int *mark = (int *)malloc(sizeof(int) * N);
    for (int i = 0; i < N; ++i) mark[i] = -1;
    /* element lists per node */
    int *ep = (int *)calloc(N + 1, sizeof(int));
    for (int e = 0; e < m->T; ++e) for (int q = 0; q < npe; ++q) ep[m->conn[e * npe + q] + 1]++;
    for (int i = 0; i < N; ++i) ep[i + 1] += ep[i];
    int *el = (int *)malloc(sizeof(int) * ep[N]);
    int *fill = (int *)malloc(sizeof(int) * N);
    memcpy(fill, ep, sizeof(int) * N);
    for (int e = 0; e < m->T; ++e) for (int q = 0; q < npe; ++q) el[fill[m->conn[e * npe + q]]++] = e;
    int nnz = 0;
    ia[0] = 0;
    for (int i = 0; i < N; ++i) {
        int start = nnz;
        for (int k = ep[i]; k < ep[i + 1]; ++k) {
            int e = el[k];
            for (int q = 0; q < npe; ++q) {
                int j = m->conn[e * npe + q];
                if (mark[j] != i) { mark[j] = i; if (ja) ja[nnz] = j; nnz++; }
            }
        }
        if (ja) { /* sort the row */
            for (int a = start + 1; a < nnz; ++a) {
                int v = ja[a], b = a - 1;
                while (b >= start && ja[b] > v) { ja[b + 1] = ja[b]; --b; }
                ja[b + 1] = v;
            }
        }
        ia[i + 1] = nnz;
    }
    free(mark); free(ep); free(el); free(fill);
    return nnz;
}

/* sequential greedy distance-2 colouring (preconditioner_dss.cpp:72-94); returns #colours */
int orc_colouring(int N, const int *ia, const int *ja, int *colour) {
    int *mark = (int *)malloc(sizeof(int) * N);
    for (int i = 0; i < N; ++i) { mark[i] = -1; colour[i] = N - 1; }
    int maxc = 0;
    for (int i = 0; i < N; ++i) {
        for (int p = ia[i]; p < ia[i + 1]; ++p) {
            int nb = ja[p];
            for (int q = ia[nb]; q < ia[nb + 1]; ++q) mark[colour[ja[q]]] = i;
        }
        int sc = 0;
        while (sc < maxc && mark[sc] == i) ++sc;
        if (sc == maxc) ++maxc;
        colour[i] = sc;
    }
    free(mark);
    return maxc;
}

/* values of -(dF/dy + c dF/dy') on the pattern (preconditioner_dss.cpp:317-357) */
void orc_fd_jacobian(OrcProblem *p, int N, double t, double c, double h, const double *y, const double *yp,
                     const double *F0, const double *ewt, const int *ia, const int *ja, const int *colour,
                     int ncol, double *vals) {
    double eps = sqrt(2.220446049250313e-16);
    double *shift = (double *)calloc(N, sizeof(double));
    double *ys = (double *)calloc(N, sizeof(double));
    double *yps = (double *)calloc(N, sizeof(double));
    double *Fs = (double *)calloc(N, sizeof(double));
    for (int j = 0; j < N; ++j) {
        double a = fabs(y[j]), b = fabs(h * yp[j]), w = 1.0 / ewt[j];
        shift[j] = eps * fmax(a, fmax(b, w));
    }
    for (int k = 0; k < ncol; ++k) {
        memcpy(ys, y, sizeof(double) * N);
        memcpy(yps, yp, sizeof(double) * N);
        for (int j = 0; j < N; ++j)
            if (colour[j] == k) { ys[j] += shift[j]; yps[j] += c * shift[j]; }
        orc_residual(p, t, ys, yps, Fs);
        for (int j = 0; j < N; ++j) {
            if (colour[j] != k) continue;
            for (int q = ia[j]; q < ia[j + 1]; ++q) {
                int row = ja[q];
                /* position of (row, j) */
                int lo = ia[row], hi = ia[row + 1] - 1, pos = -1;
                while (lo <= hi) {
                    int mid = (lo + hi) / 2;
                    if (ja[mid] == j) { pos = mid; break; }
                    if (ja[mid] < j) lo = mid + 1; else hi = mid - 1;
                }
                vals[pos] = (F0[row] - Fs[row]) / shift[j];
            }
        }
    }
    free(shift); free(ys); free(yps); free(Fs);
}

/* FD J v with the increment of the product: sqrt(eps) ||y|| / ||v|| */
typedef struct {
    OrcProblem *p; int N; double t, c; const double *y, *yp, *F0;
    double *ys, *yps, *F1;
} JvOp;
static void jv_op(void *ctx, const double *v, double *out) {
    JvOp *o = (JvOp *)ctx;
    double ny = dnrm2(o->N, o->y), nv = dnrm2(o->N, v);
    if (ny == 0.) ny = 1.;
    double sig = nv > 0 ? sqrt(2.220446049250313e-16) * ny / nv : 1.;
    for (int i = 0; i < o->N; ++i) { o->ys[i] = o->y[i] + sig * v[i]; o->yps[i] = o->yp[i] + o->c * (sig * v[i]); }
    orc_residual(o->p, o->t, o->ys, o->yps, o->F1);
    for (int i = 0; i < o->N; ++i) out[i] = (o->F1[i] - o->F0[i]) / sig;
}
void orc_jv(OrcProblem *p, int N, double t, const double *y, const double *yp, const double *F0, double c,
            const double *v, double *out) {
    JvOp o = {p, N, t, c, y, yp, F0, NULL, NULL, NULL};
    o.ys = (double *)calloc(N, sizeof(double));
    o.yps = (double *)calloc(N, sizeof(double));
    o.F1 = (double *)calloc(N, sizeof(double));
    jv_op(&o, v, out);
    free(o.ys); free(o.yps); free(o.F1);
}

/* Newton linear solve J x = b, matrix-free J, preconditioner ILU(0)/Jacobi of the
 * FD Jacobian values (pattern ia, ja).  Returns 1 on success. */
int orc_newton_gmres(OrcProblem *p, int N, double t, const double *y, const double *yp, const double *F0, double c,
                     const int *ia, const int *ja, const double *jvals, int precon_type, const double *b, double *x,
                     const int *iparams, double tol, double *stats) {
    JvOp o = {p, N, t, c, y, yp, F0, NULL, NULL, NULL};
    o.ys = (double *)calloc(N, sizeof(double));
    o.yps = (double *)calloc(N, sizeof(double));
    o.F1 = (double *)calloc(N, sizeof(double));
    PcOp P = {N, precon_type, ia, ja, NULL, NULL, NULL};
    int ok = 1;
    if (precon_type == 5) {
        P.lu = (double *)calloc(ia[N], sizeof(double));
        memcpy(P.lu, jvals, sizeof(double) * ia[N]);
        P.diagpos = (int *)malloc(sizeof(int) * N);
        ok = orc_ilu0_factor(N, ia, ja, P.lu, P.diagpos);
    } else if (precon_type == 2) {
        P.J = (double *)calloc(N, sizeof(double));
        ok = orc_jacobi_factor(N, ia, ja, jvals, P.J);
    }
    if (ok) {
        OrcGmres run = {iparams[0], iparams[1], iparams[2], tol, 0, 0, 0, 0., 0.};
        orc_gmres(N, jv_op, &o, precon_type ? pc_op : NULL, &P, b, x, &run);
        stats[0] = run.restarts; stats[1] = run.total_iters; stats[2] = run.converged;
        stats[3] = run.residual; stats[4] = run.error;
    }
    free(P.lu); free(P.diagpos); free(P.J); free(o.ys); free(o.yps); free(o.F1);
    return ok;
}
