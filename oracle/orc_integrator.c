/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle of the time integration: the
 * variable-step BDF corrector (Lagrange-form coefficients), the predictor
 * through the last k+1 solution values and the Newton-GMRES corrector solve,
 * replayed over a given step sequence (step sizes, orders and the frozen
 * upwind weights of every step).  The reference obtains these from SUNDIALS
 * IDAS (ida_integrator.h:169-340, 296-321), which the reference does not ship:
 * the published method (Brenan, Campbell & Petzold, ch. 3) is restated here
 * and parity of the adaptive control itself is checked through the replayed
 * sequence plus physical invariants (mass balance), not against IDAS.
 * Newton is iterated to a tight tolerance so the result is the BDF solution
 * of each step.  Consistent initial derivative: Physics::initialise
 * (fvmpor_ODE_impl.cpp:34-60).
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct OrcProblem OrcProblem;
void orc_residual(OrcProblem *p, double t, const double *h, const double *hp, double *res);
void orc_set_weights(OrcProblem *p, const double *wf);
void orc_get_node_fields(const OrcProblem *p, double *ahh, double *Sw, double *dSw);
int orc_num_dirichlet(const OrcProblem *p);
void orc_dirichlet(const OrcProblem *p, int *nodes, double *vals);
int orc_jac_pattern(const OrcMesh *m, int *ia, int *ja);
int orc_colouring(int N, const int *ia, const int *ja, int *colour);
void orc_fd_jacobian(OrcProblem *p, int N, double t, double c, double h, const double *y, const double *yp,
                     const double *F0, const double *ewt, const int *ia, const int *ja, const int *colour,
                     int ncol, double *vals);
int orc_ilu0_factor(int n, const int *ia, const int *ja, double *nz, int *diagpos);
void orc_ilu0_apply(int n, const int *ia, const int *ja, const double *nz, const int *diagpos, const double *x,
                    double *y);
typedef void (*OpFn)(void *ctx, const double *x, double *y);
typedef struct {
    int dim_krylov, max_restarts, reorthogonalise;
    double tol;
    int restarts, total_iters, converged;
    double residual, error;
} OrcGmres;
void orc_gmres(int n, OpFn A, void *actx, OpFn M, void *mctx, const double *b, double *x, OrcGmres *run);
OrcMesh *orc_problem_mesh(OrcProblem *p);

/* consistent psi'(t0) = F / ahh (psi' = 0 on Dirichlet rows and where ahh = 0) */
void orc_initial_derivative(OrcProblem *p, int N, double t0, const double *y0, double *yp0) {
    double *F = (double *)calloc(N, sizeof(double));
    double *ahh = (double *)calloc(N, sizeof(double));
    memset(yp0, 0, sizeof(double) * N);
    orc_residual(p, t0, y0, yp0, F);
    orc_get_node_fields(p, ahh, NULL, NULL);
    for (int i = 0; i < N; ++i) yp0[i] = ahh[i] != 0. ? F[i] / ahh[i] : 0.;
    int nd = orc_num_dirichlet(p);
    if (nd) {
        int *dn = (int *)malloc(sizeof(int) * nd);
        double *dv = (double *)malloc(sizeof(double) * nd);
        orc_dirichlet(p, dn, dv);
        for (int k = 0; k < nd; ++k) yp0[dn[k]] = 0.;
        free(dn); free(dv);
    }
    free(F); free(ahh);
}

typedef struct {
    OrcProblem *p; int N; double t, c; const double *y, *yp, *F0, *ewt;
    double *ys, *yps, *F1;
} SJvOp;
/* (S J) v with S = diag(ewt), J v by the FD quotient of the product */
static void sjv(void *ctx, const double *v, double *out) {
    SJvOp *o = (SJvOp *)ctx;
    double ny = 0., nv = 0.;
    for (int i = 0; i < o->N; ++i) { ny += o->y[i] * o->y[i]; nv += v[i] * v[i]; }
    ny = sqrt(ny); nv = sqrt(nv);
    if (ny == 0.) ny = 1.;
    double sig = nv > 0 ? sqrt(2.220446049250313e-16) * ny / nv : 1.;
    for (int i = 0; i < o->N; ++i) { o->ys[i] = o->y[i] + sig * v[i]; o->yps[i] = o->yp[i] + o->c * (sig * v[i]); }
    orc_residual(o->p, o->t, o->ys, o->yps, o->F1);
    for (int i = 0; i < o->N; ++i) out[i] = (o->F1[i] - o->F0[i]) / sig * o->ewt[i];
}
typedef struct { int n; const int *ia, *ja; double *lu; int *dp; const double *ewt; double *tmp; } SPc;
static void spc(void *ctx, const double *x, double *y) {
    SPc *P = (SPc *)ctx;
    for (int i = 0; i < P->n; ++i) P->tmp[i] = x[i] / P->ewt[i];
    orc_ilu0_apply(P->n, P->ia, P->ja, P->lu, P->dp, P->tmp, y);
}

static void pred_weights(const double *tau, int q, double *w) {
    for (int i = 1; i <= q + 1; ++i) {
        double v = 1.;
        for (int j = 1; j <= q + 1; ++j) if (j != i) v *= (tau[0] - tau[j]) / (tau[i] - tau[j]);
        w[i - 1] = v;
    }
}
static void bdf_weights(const double *tau, int k, double *a) {
    double s = 0.;
    for (int j = 1; j <= k; ++j) s += 1. / (tau[0] - tau[j]);
    a[0] = s;
    for (int i = 1; i <= k; ++i) {
        double num = 1., den = 1.;
        for (int j = 1; j <= k; ++j) if (j != i) num *= (tau[0] - tau[j]);
        for (int j = 0; j <= k; ++j) if (j != i) den *= (tau[i] - tau[j]);
        a[i] = num / den;
    }
}

/* Replay nsteps steps of sizes hs[] and orders ks[], upwind front weights
 * wfs[step * NE ...] per step.  rtol/atol define the weights of the scaled
 * linear solves and the Newton stopping test ||delta||_WRMS <= newton_tol.
 * Returns the number of Newton iterations, or -1 on a failed step. */
int orc_integrate_replay(OrcProblem *p, double rtol, double atol, double newton_tol, double t0, const double *psi0,
                         int nsteps, const double *hs, const int *ks, const double *wfs, int NE, double *psi_out) {
    OrcMesh *m = orc_problem_mesh(p);
    const int N = m->N;
    enum { NH = 7 };
    double *hist[NH], th[NH];
    for (int i = 0; i < NH; ++i) { hist[i] = (double *)calloc(N, sizeof(double)); th[i] = t0; }
    memcpy(hist[0], psi0, sizeof(double) * N);
    int nh = 1;
    double *yp = (double *)calloc(N, sizeof(double));
    double *y = (double *)calloc(N, sizeof(double));
    double *ypn = (double *)calloc(N, sizeof(double));
    double *beta = (double *)calloc(N, sizeof(double));
    double *G = (double *)calloc(N, sizeof(double));
    double *b = (double *)calloc(N, sizeof(double));
    double *dl = (double *)calloc(N, sizeof(double));
    double *ewt = (double *)calloc(N, sizeof(double));
    double *tmp = (double *)calloc(N, sizeof(double));
    int *ia = (int *)calloc(N + 1, sizeof(int));
    int nnz = orc_jac_pattern(m, ia, NULL);
    int *ja = (int *)calloc(nnz, sizeof(int));
    orc_jac_pattern(m, ia, ja);
    int *col = (int *)calloc(N, sizeof(int));
    int ncol = orc_colouring(N, ia, ja, col);
    double *J = (double *)calloc(nnz, sizeof(double));
    int *dp = (int *)calloc(N, sizeof(int));
    SJvOp op = {p, N, 0, 0, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
    op.ys = (double *)calloc(N, sizeof(double));
    op.yps = (double *)calloc(N, sizeof(double));
    op.F1 = (double *)calloc(N, sizeof(double));
    int newton_total = 0, status = 0;

    double *half = (double *)malloc(sizeof(double) * NE);
    for (int e = 0; e < NE; ++e) half[e] = 0.5;
    orc_set_weights(p, half);
    orc_initial_derivative(p, N, t0, hist[0], yp);
    free(half);
    for (int st = 0; st < nsteps && !status; ++st) {
        double h = hs[st];
        int k = ks[st];
        orc_set_weights(p, wfs + (size_t)st * NE);
        for (int i = 0; i < N; ++i) ewt[i] = 1. / (rtol * fabs(hist[0][i]) + atol);
        double tau[NH + 1];
        tau[0] = th[0] + h;
        for (int i = 0; i < nh; ++i) tau[i + 1] = th[i];
        int q = k < nh - 1 ? k : nh - 1;
        if (q < 1) {
            for (int i = 0; i < N; ++i) y[i] = hist[0][i] + h * yp[i];
        } else {
            double w[NH];
            pred_weights(tau, q, w);
            for (int i = 0; i < N; ++i) {
                double s = 0.;
                for (int r = 0; r <= q; ++r) s += w[r] * hist[r][i];
                y[i] = s;
            }
        }
        double a[NH];
        bdf_weights(tau, k, a);
        for (int i = 0; i < N; ++i) {
            double s = 0.;
            for (int r = 1; r <= k; ++r) s += a[r] * hist[r - 1][i];
            beta[i] = s;
        }
        const double cj = a[0];
        int conv = 0;
        for (int it = 0; it < 30; ++it) {
            for (int i = 0; i < N; ++i) ypn[i] = cj * y[i] + beta[i];
            orc_residual(p, tau[0], y, ypn, G);
            orc_fd_jacobian(p, N, tau[0], cj, h, y, ypn, G, ewt, ia, ja, col, ncol, J);
            if (!orc_ilu0_factor(N, ia, ja, J, dp)) { status = -1; break; }
            for (int i = 0; i < N; ++i) { b[i] = -G[i] * ewt[i]; dl[i] = 0.; }
            op.t = tau[0]; op.c = cj; op.y = y; op.yp = ypn; op.F0 = G; op.ewt = ewt;
            SPc P = {N, ia, ja, J, dp, ewt, tmp};
            OrcGmres run = {30, 20, 1, 1e-12, 0, 0, 0, 0., 0.};
            orc_gmres(N, sjv, &op, spc, &P, b, dl, &run);
            double dn = 0.;
            for (int i = 0; i < N; ++i) { y[i] += dl[i]; dn += (dl[i] * ewt[i]) * (dl[i] * ewt[i]); }
            dn = sqrt(dn / N);
            ++newton_total;
            if (dn <= newton_tol) { conv = 1; break; }
        }
        if (!conv) { status = -1; break; }
        for (int i = 0; i < N; ++i) ypn[i] = cj * y[i] + beta[i];
        double *old = hist[NH - 1];
        for (int r = NH - 1; r > 0; --r) { hist[r] = hist[r - 1]; th[r] = th[r - 1]; }
        hist[0] = old;
        th[0] = tau[0];
        memcpy(hist[0], y, sizeof(double) * N);
        memcpy(yp, ypn, sizeof(double) * N);
        if (nh < NH) ++nh;
    }
    memcpy(psi_out, hist[0], sizeof(double) * N);
    for (int i = 0; i < NH; ++i) free(hist[i]);
    free(yp); free(y); free(ypn); free(beta); free(G); free(b); free(dl); free(ewt); free(tmp); free(ia); free(ja);
    free(col); free(J); free(dp); free(op.ys); free(op.yps); free(op.F1);
    return status ? -1 : newton_total;
}
