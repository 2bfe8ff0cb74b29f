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
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct OrcProblem OrcProblem;
void orc_residual(OrcProblem *p, double t, const double *h, const double *hp, double *res);
void orc_set_weights(OrcProblem *p, const double *wf);
void orc_begin_step(OrcProblem *p, double t, const double *psi, const double *psip);
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

/* Options of a replay.
 *  newton_tol   Newton stopping test ||delta||_WRMS <= newton_tol (ida_test = 0),
 *               or IDA's rate-based test ss * ||delta|| <= newton_tol (ida_test = 1,
 *               ss = 20 after a new iteration matrix, then rate / (1 - rate)).
 *  lin_rel      GMRES relative tolerance (lin_abs = 0), or
 *  lin_abs      IDA's linear tolerance lin_abs * sqrt(N) on the scaled residual
 *               (eplifac * epsNewt; IDASpgmr(ida_mem, 0) at ida_integrator.h:252
 *               keeps IDA's defaults 0.05 and 0.33).
 *  jac_policy   0: coloured FD Jacobian + ILU(0) at every Newton iteration;
 *               1: once per step; 2: IDA's lsetup policy (first step, c_j moved
 *               by more than 25 %, after a failed Newton iteration, every 20 steps).
 *  max_newton, dim_krylov, max_restarts: iteration limits.
 *  own_upwind   1: freeze the upwind weights of every step from the state at its
 *               start (preprocess_timestep, fvmpor_ODE_impl.cpp:118-142) instead
 *               of taking them from wfs.
 *  verbose      1: one progress line per Newton iteration on stderr. */
typedef struct {
    double newton_tol, lin_rel, lin_abs;
    int jac_policy, ida_test, max_newton, dim_krylov, max_restarts, own_upwind, verbose;
} OrcReplayOpts;

/* Replay nsteps steps of sizes hs[] and orders ks[], upwind front weights
 * wfs[step * NE ...] per step.  rtol/atol define the weights of the scaled
 * linear solves.  Returns the number of Newton iterations (stats, if not
 * null, receives Newton iterations, GMRES iterations, Jacobian set-ups and
 * residual evaluations), or -1 on a failed step. */
int orc_integrate_replay_opts(OrcProblem *p, double rtol, double atol, const OrcReplayOpts *o, double t0,
                              const double *psi0, int nsteps, const double *hs, const int *ks, const double *wfs,
                              int NE, double *psi_out, double *stats);

int orc_integrate_replay(OrcProblem *p, double rtol, double atol, double newton_tol, double t0, const double *psi0,
                         int nsteps, const double *hs, const int *ks, const double *wfs, int NE, double *psi_out) {
    OrcReplayOpts o = {newton_tol, 1e-12, 0., 0, 0, 30, 30, 20, 0, 0};
    return orc_integrate_replay_opts(p, rtol, atol, &o, t0, psi0, nsteps, hs, ks, wfs, NE, psi_out, NULL);
}

int orc_integrate_replay_opts(OrcProblem *p, double rtol, double atol, const OrcReplayOpts *o, double t0,
                              const double *psi0, int nsteps, const double *hs, const int *ks, const double *wfs,
                              int NE, double *psi_out, double *stats) {
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
    long n_gmres = 0, n_setup = 0, n_res = 0;
    double cj_setup = -1., ss = 20.;
    int since_setup = 0, need_setup = 1;
    for (int st = 0; st < nsteps && !status; ++st) {
        double h = hs[st];
        /* an order above the number of stored solutions (a replay that starts
         * inside a run) falls back to the highest order the history allows */
        int k = ks[st] < nh ? ks[st] : nh;
        if (o->own_upwind) orc_begin_step(p, th[0], hist[0], yp);
        else orc_set_weights(p, wfs + (size_t)st * NE);
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
        double oldnrm = 0.;
        for (int it = 0; it < o->max_newton; ++it) {
            for (int i = 0; i < N; ++i) ypn[i] = cj * y[i] + beta[i];
            orc_residual(p, tau[0], y, ypn, G);
            ++n_res;
            const double ratio = cj_setup > 0 ? cj / cj_setup : 0.;
            int setup = o->jac_policy == 0 || (o->jac_policy == 1 && it == 0) ||
                        (o->jac_policy == 2 && it == 0 &&
                         (need_setup || ratio < 0.75 || ratio > 1.25 || since_setup >= 20));
            if (setup) {
                orc_fd_jacobian(p, N, tau[0], cj, h, y, ypn, G, ewt, ia, ja, col, ncol, J);
                n_res += ncol;
                ++n_setup;
                if (!orc_ilu0_factor(N, ia, ja, J, dp)) { status = -1; break; }
                cj_setup = cj;
                need_setup = 0;
                since_setup = 0;
                ss = 20.;
            }
            double bn = 0.;
            for (int i = 0; i < N; ++i) { b[i] = -G[i] * ewt[i]; dl[i] = 0.; bn += b[i] * b[i]; }
            bn = sqrt(bn);
            op.t = tau[0]; op.c = cj; op.y = y; op.yp = ypn; op.F0 = G; op.ewt = ewt;
            SPc P = {N, ia, ja, J, dp, ewt, tmp};
            double tol = o->lin_rel;
            if (o->lin_abs > 0.) tol = bn > 0. ? o->lin_abs * sqrt((double)N) / bn : 1.;
            OrcGmres run = {o->dim_krylov, o->max_restarts, 1, tol, 0, 0, 0, 0., 0.};
            orc_gmres(N, sjv, &op, spc, &P, b, dl, &run);
            n_gmres += run.total_iters;
            n_res += run.total_iters;
            double dn = 0.;
            for (int i = 0; i < N; ++i) { y[i] += dl[i]; dn += (dl[i] * ewt[i]) * (dl[i] * ewt[i]); }
            dn = sqrt(dn / N);
            ++newton_total;
            if (o->verbose) {
                fprintf(stderr, "oracle replay: step %d Newton %d: %d GMRES iterations, |delta| %.3e\n", st, it,
                        run.total_iters, dn);
                fflush(stderr);
            }
            if (!o->ida_test) {
                if (dn <= o->newton_tol) { conv = 1; break; }
            } else if (it == 0) {
                oldnrm = dn;
                if (dn <= 1e-4 * o->newton_tol || ss * dn <= o->newton_tol) { conv = 1; break; }
            } else {
                double rate = pow(dn / oldnrm, 1.0 / it);
                if (rate > 0.9) break;
                ss = fmax(rate / (1. - rate), 0.05);
                if (ss * dn <= o->newton_tol) { conv = 1; break; }
            }
        }
        if (!conv) { status = -1; break; }
        ++since_setup;
        for (int i = 0; i < N; ++i) ypn[i] = cj * y[i] + beta[i];
        double *old = hist[NH - 1];
        for (int r = NH - 1; r > 0; --r) { hist[r] = hist[r - 1]; th[r] = th[r - 1]; }
        hist[0] = old;
        th[0] = tau[0];
        memcpy(hist[0], y, sizeof(double) * N);
        memcpy(yp, ypn, sizeof(double) * N);
        if (nh < NH) ++nh;
    }
    if (stats) {
        stats[0] = newton_total; stats[1] = (double)n_gmres; stats[2] = (double)n_setup; stats[3] = (double)n_res;
    }
    memcpy(psi_out, hist[0], sizeof(double) * N);
    for (int i = 0; i < NH; ++i) free(hist[i]);
    free(yp); free(y); free(ypn); free(beta); free(G); free(b); free(dl); free(ewt); free(tmp); free(ia); free(ja);
    free(col); free(J); free(dp); free(op.ys); free(op.yps); free(op.F1);
    return status ? -1 : newton_total;
}
