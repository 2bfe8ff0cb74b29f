// Variable-step, variable-order BDF integrator for F(t, psi, psi') = 0.
//
// The reference delegates time stepping to SUNDIALS IDAS (IDAIntegrator,
// ida_integrator.h:169-340); IDAS is not part of the reference, so this is an
// implementation of the published method IDAS uses (Brenan, Campbell & Petzold,
// "Numerical Solution of Initial-Value Problems in DAEs", ch. 3; Hindmarsh et al.,
// "SUNDIALS", ACM TOMS 31 (2005)): BDF of order 1..max_order in variable-step
// Lagrange form, a predictor through the last k+1 solution values, modified
// Newton on G(y) = F(t, y, alpha_0 y + beta), linear solves by GMRES with the
// IDA scalings (left by the error weights, right-preconditioned by the coloured
// FD Jacobian's ILU(0) or Jacobi times the inverse weights), the IDA Newton
// rate test and linear tolerance delta = eplifac * epsNewt * sqrt(N), and a
// local error estimate from the predictor-corrector difference.
//
// Host code here only computes scalar coefficients; every vector operation
// runs on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fvp_blas.h"
#include "fvp_internal.h"

using fvp::DevBuf;
using fvp::set_error;

namespace {
constexpr int KMAX = 5;
constexpr int NHIST = KMAX + 2;
}

struct fvp_integrator {
    fvp_ctx *c = nullptr;
    fvp_integrator_params p{};
    int N = 0;
    // history: hist[slot[0]] = y at th[0] = t_n, hist[slot[1]] at th[1], ...
    DevBuf<double> hist[NHIST];
    int slot[NHIST];
    double th[NHIST];
    int nh = 0;
    DevBuf<double> yp, ewt, y, ypn, ypred, ypred2, G, b, delta, beta, tmp, part, scal;
    double *hscal = nullptr;   // pinned host: the norms the step logic reads
    hipEvent_t hev = nullptr;  // marks the kernels that wrote them
    double h = 0., ss = 20., cj_setup = -1.;
    int k = 1, steps_at_k = 0, steps_since_setup = 0;
    bool need_setup = true;
    long nres0 = 0;
    fvp_integrator_stats st{};
    // independent instances stored back to back (an ensemble): norms are the
    // maximum over instances of the per-instance weighted RMS
    int nseg = 0, seg_min = 0;
    DevBuf<int> seg;
    DevBuf<double> seg_sum;
    // learned initial iterate (network, its features / output, the candidate
    // state) and training-pair recording into caller buffers
    fvp_mpnn *mp = nullptr;
    DevBuf<float> stat, feat, ylearn;  // static feature columns, features, network output
    DevBuf<double> y2, ypn2, G2;
    float *rec_feat = nullptr, *rec_targ = nullptr;
    double learn_trust = 2., learn_band = 0.5;
    bool learn_took = false;   // this attempt started Newton from the network's proposal
    bool learn_block = false;  // retries of this step start from the predictor
    // back-off: after kLearnWeakRun tries in a row whose candidate did not at
    // least halve the preconditioned residual, the network sits out learn_skip
    // attempts (4, 8, ... up to 64; a strong try resets it)
    int learn_weak = 0, learn_skip = 0, learn_backoff = 4;
};

namespace fvp {

// learned start: a try is strong when the candidate's preconditioned residual
// is below this share of the predictor's; weak tries in a row before a back-off
// (one: a weak try costs a network pass, a residual and two preconditioner
// solves, and on the 1M-node box no try was strong)
constexpr double kLearnStrong = 0.5;
constexpr int kLearnWeakRun = 1;

struct Coef8 {
    double v[8];
    const double *x[8];
};

// out = sum_i c_i x_i (up to 8 vectors)
__global__ void lincomb8(long n, int m, Coef8 cf, double *__restrict__ out) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.;
    for (int q = 0; q < m; ++q) s += cf.v[q] * cf.x[q][i];
    out[i] = s;
}

__global__ void ewt_kernel(long n, double rtol, double atol, const double *__restrict__ y, double *__restrict__ w) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) w[i] = 1.0 / (rtol * fabs(y[i]) + atol);
}

__global__ void mul_kernel(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ o) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] * b[i];
}

__global__ void div_kernel(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ o) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] / b[i];
}

__global__ void axpby_kernel(long n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// yp0 = F / ahh where ahh != 0 and the row is not Dirichlet, else 0 (Physics::initialise)
__global__ void ic_deriv(long n, const double *__restrict__ F, const double *__restrict__ ahh,
                         double *__restrict__ yp) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) yp[i] = ahh[i] != 0. ? F[i] / ahh[i] : 0.;
}

__global__ void zero_rows(int nd, const int *__restrict__ rows, double *__restrict__ v) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nd) v[rows[k]] = 0.;
}

static inline int nb(long n) { return (int)((n + 255) / 256); }

// node features of the learned iterate: columns 0..4 from the step, 5..15 static.
// Column 4 is the Newton correction of the capacity-diagonal approximation of
// the iteration matrix (dF/dpsi' = -ahh dominates for small h), in error-weight
// units: ewt G / (c_j ahh), 0 where ahh vanishes.
__global__ void learn_features(long n, const double *__restrict__ yn, const double *__restrict__ ypred,
                               const double *__restrict__ ewt, const double *__restrict__ G,
                               const double *__restrict__ ahh, double cj, float lh, float kk,
                               const float *__restrict__ stat, float *__restrict__ x) {
    // four lanes per node, one 16-byte column group each
    const long g = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long i = g >> 2;
    const int q = (int)(g & 3);
    if (i >= n) return;
    float4 o = stat ? reinterpret_cast<const float4 *>(stat)[g] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (q == 0) {
        o.x = (float)yn[i];
        o.y = (float)asinh((ypred[i] - yn[i]) * ewt[i]);
        o.z = lh;
        o.w = kk;
    } else if (q == 1) {
        const double d = cj * ahh[i];
        o.x = d > 0. ? (float)asinh(ewt[i] * G[i] / d) : 0.f;
    }
    reinterpret_cast<float4 *>(x)[g] = o;
}

// learned candidate y2 = ypred + c / ewt from the network output dy (error-weight
// units), with c limited by the capacity-diagonal Newton step
// jac = ewt G / (c_j ahh) of the node (0 where ahh = 0):
//   |c| <= trust |jac| and c of the sign of jac -- a node whose equation the
//   predictor already satisfies is not moved;
//   c = 0 where ypred > -band (the near-saturated band, where the van
//   Genuchten-Mualem conductivity's derivative is unbounded and a stray
//   correction dominates the next residual and the linear solve)
__global__ void learn_guess(long n, const double *__restrict__ ypred, const float *__restrict__ dy,
                            const double *__restrict__ ewt, const double *__restrict__ G,
                            const double *__restrict__ ahh, double cj, double trust, double band,
                            double *__restrict__ y2) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = cj * ahh[i];
    const double jac = d > 0. ? ewt[i] * G[i] / d : 0.;
    const double lim = trust * std::fabs(jac);
    double c = std::fmin(std::fmax((double)dy[i], -lim), lim);
    if (c * jac < 0. || ypred[i] > -band) c = 0.;
    y2[i] = ypred[i] + c / ewt[i];
}

// training target ewt * (y - ypred)
__global__ void learn_target(long n, const double *__restrict__ y, const double *__restrict__ ypred,
                             const double *__restrict__ ewt, float *__restrict__ t) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) t[i] = (float)((y[i] - ypred[i]) * ewt[i]);
}

__global__ void zero_rows_f(int nd, const int *__restrict__ rows, float *__restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd) v[rows[i]] = 0.f;
}

// sum over each instance's nodes of (v ewt)^2: one workgroup per instance
__global__ void __launch_bounds__(256) seg_sq(const int *__restrict__ seg, const double *__restrict__ v,
                                              const double *__restrict__ ewt, double *__restrict__ out) {
    __shared__ double lds[4];
    const int g = blockIdx.x;
    double acc = 0.;
    for (int i = seg[g] + threadIdx.x; i < seg[g + 1]; i += blockDim.x) {
        const double x = v[i] * ewt[i];
        acc += x * x;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[g] = lds[0] + lds[1] + lds[2] + lds[3];
}

// out[0] = max over instances of sqrt(sum / size), out[1] = total sum; the
// same to hout (pinned host memory)
__global__ void seg_fold(int nseg, const int *__restrict__ seg, const double *__restrict__ sums,
                         double *__restrict__ out, double *__restrict__ hout) {
    if (threadIdx.x != 0) return;
    double mx = 0., tot = 0.;
    for (int g = 0; g < nseg; ++g) {
        mx = fmax(mx, sqrt(sums[g] / (seg[g + 1] - seg[g])));
        tot += sums[g];
    }
    out[0] = mx;
    out[1] = tot;
    hout[0] = mx;
    hout[1] = tot;
    __threadfence_system();
}

// Weighted RMS norm ||v||_WRMS with the error weights (the maximum over the
// instances of an ensemble); l2 (optional) gets the plain 2-norm of v * ewt.
// The norm's last kernel stores it to pinned host memory; a failed wait gives
// NaN (every test on it fails; the error is recorded).
static double wrms(fvp_integrator *I, const double *v, double *l2 = nullptr) {
    hipStream_t s = I->c->stream;
    const double *r = I->hscal;
    if (I->nseg > 1) {
        seg_sq<<<I->nseg, 256, 0, s>>>(I->seg.p, v, I->ewt.p, I->seg_sum.p);
        seg_fold<<<1, 64, 0, s>>>(I->nseg, I->seg.p, I->seg_sum.p, I->scal.p, I->hscal);
        if (host_wait(s, I->hev)) return std::nan("");
        if (l2) *l2 = std::sqrt(r[1]);
        return r[0];
    }
    mul_kernel<<<nb(I->N), 256, 0, s>>>(I->N, v, I->ewt.p, I->tmp.p);
    dot_async(s, I->N, I->tmp.p, I->tmp.p, I->part.p, I->scal.p, I->hscal);
    if (host_wait(s, I->hev)) return std::nan("");
    if (l2) *l2 = std::sqrt(r[0]);
    return std::sqrt(r[0] / I->N);
}

static void lincomb(fvp_integrator *I, int m, const double *c, const double *const *x, double *out) {
    Coef8 cf{};
    for (int q = 0; q < m; ++q) { cf.v[q] = c[q]; cf.x[q] = x[q]; }
    lincomb8<<<nb(I->N), 256, 0, I->c->stream>>>(I->N, m, cf, out);
}

// Lagrange value weights at tau[0] of the polynomial through tau[1..q+1]
static void predictor_weights(const double *tau, int q, double *w) {
    for (int i = 1; i <= q + 1; ++i) {
        double v = 1.;
        for (int j = 1; j <= q + 1; ++j)
            if (j != i) v *= (tau[0] - tau[j]) / (tau[i] - tau[j]);
        w[i - 1] = v;
    }
}

// BDF derivative weights at tau[0] over nodes tau[0..k]: y'(tau0) ~ sum_i a_i y(tau_i)
static void bdf_weights(const double *tau, int k, double *a) {
    double s = 0.;
    for (int j = 1; j <= k; ++j) s += 1. / (tau[0] - tau[j]);
    a[0] = s;
    for (int i = 1; i <= k; ++i) {
        double num = 1., den = 1.;
        for (int j = 1; j <= k; ++j)
            if (j != i) num *= (tau[0] - tau[j]);
        for (int j = 0; j <= k; ++j)
            if (j != i) den *= (tau[i] - tau[j]);
        a[i] = num / den;
    }
}

static const double *H(fvp_integrator *I, int i) { return I->hist[I->slot[i]].p; }

// predicted y at tau[0] of degree q (needs q+1 history points); q < 0: y_n + h y'_n
static void predict(fvp_integrator *I, const double *tau, int q, double *out) {
    if (q < 1 || I->nh < q + 1) {
        double c2[2] = {1., tau[0] - tau[1]};
        const double *x2[2] = {H(I, 0), I->yp.p};
        lincomb(I, 2, c2, x2, out);
        return;
    }
    double w[NHIST];
    const double *x[NHIST];
    predictor_weights(tau, q, w);
    for (int i = 0; i <= q; ++i) x[i] = H(I, i);
    lincomb(I, q + 1, w, x, out);
}

// One attempt: 0 accepted, 1 error test failed, 2 Newton / linear failure, <0 fatal
static int attempt(fvp_integrator *I, double h, int k, double *err_k, double *err_km1) {
    fvp_ctx *c = I->c;
    hipStream_t s = c->stream;
    const long N = I->N;
    double tau[NHIST + 1];
    tau[0] = I->th[0] + h;
    for (int i = 0; i < I->nh; ++i) tau[i + 1] = I->th[i];
    const double t = tau[0];
    const int q = std::min(k, I->nh - 1);  // predictor degree
    predict(I, tau, q, I->ypred.p);
    double a[NHIST];
    bdf_weights(tau, k, a);
    const double cj = a[0];
    // beta = sum_{i>=1} a_i y_{n+1-i}
    {
        double w[NHIST];
        const double *x[NHIST];
        for (int i = 1; i <= k; ++i) { w[i - 1] = a[i]; x[i - 1] = H(I, i - 1); }
        lincomb(I, k, w, x, I->beta.p);
    }
    FVP_HIP(hipMemcpyAsync(I->y.p, I->ypred.p, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    // y' = cj y + beta
    auto deriv = [&]() {
        double c2[2] = {cj, 1.};
        const double *x2[2] = {I->y.p, I->beta.p};
        lincomb(I, 2, c2, x2, I->ypn.p);
    };
    deriv();
    if (fvp::residual_impl(c, t, I->y.p, I->ypn.p, I->G.p, c->averaging)) return -1;
    I->learn_took = false;
    bool learn = I->mp && !I->learn_block;
    if (learn && I->learn_skip > 0) {
        --I->learn_skip;
        learn = false;
    }
    if (learn || I->rec_feat) {
        float *x = I->rec_feat ? I->rec_feat : I->feat.p;
        learn_features<<<nb(4 * N), 256, 0, s>>>(N, H(I, 0), I->ypred.p, I->ewt.p, I->G.p, c->cap.p, cj,
                                             (float)std::log10(h), k / 3.f, I->stat.p, x);
        FVP_HIP(hipGetLastError());
        if (learn) {
            // learned candidate and its first Newton residual
            if (fvp_mpnn_forward(I->mp, x, I->ylearn.p)) return -1;
            if (c->ndir) zero_rows_f<<<(c->ndir + 255) / 256, 256, 0, s>>>(c->ndir, c->dnode.p, I->ylearn.p);
            learn_guess<<<nb(N), 256, 0, s>>>(N, I->ypred.p, I->ylearn.p, I->ewt.p, I->G.p, c->cap.p, cj,
                                              I->learn_trust, I->learn_band, I->y2.p);
            double c2[2] = {cj, 1.};
            const double *x2[2] = {I->y2.p, I->beta.p};
            lincomb(I, 2, c2, x2, I->ypn2.p);
            if (fvp::residual_impl(c, t, I->y2.p, I->ypn2.p, I->G2.p, c->averaging)) return -1;
        }
    }
    // preconditioner setup when stale (IDA: first step, after a failure, c_j moved by > 25%)
    const double cjratio = I->cj_setup > 0 ? cj / I->cj_setup : 0.;
    if (I->p.precon_type != FVP_PRECON_NONE &&
        (I->need_setup || cjratio < 0.75 || cjratio > 1.25 || I->steps_since_setup >= 20)) {
        int rc = fvp_precon_setup(c, t, cj, h, I->y.p, I->ypn.p, I->G.p, I->ewt.p, I->p.precon_type);
        if (rc < 0) return -1;
        if (rc > 0) return 2;
        ++I->st.precon_setups;
        I->cj_setup = cj;
        I->need_setup = false;
        I->steps_since_setup = 0;
        I->ss = 20.;
    }
    if (learn) {
        // judge the learned candidate against the predictor by the size of the
        // Newton step each would start with, the preconditioned residual
        // P G (P ~ J^-1) in the error-weighted RMS norm; without a
        // preconditioner by the weighted residual itself
        double g0, g1;
        if (I->p.precon_type != FVP_PRECON_NONE) {
            if (fvp_precon_solve(c, I->G.p, I->b.p) || fvp_precon_solve(c, I->G2.p, I->delta.p)) return -1;
            g0 = wrms(I, I->b.p);
            g1 = wrms(I, I->delta.p);
        } else {
            g0 = wrms(I, I->G.p);
            g1 = wrms(I, I->G2.p);
        }
        ++I->st.learned_tried;
        if (g1 < g0) {
            ++I->st.learned_used;
            I->learn_took = true;
            std::swap(I->y, I->y2);
            std::swap(I->ypn, I->ypn2);
            std::swap(I->G, I->G2);
        }
        if (g1 < kLearnStrong * g0) {
            I->learn_weak = 0;
            I->learn_backoff = 4;
        } else if (++I->learn_weak >= kLearnWeakRun) {
            I->learn_skip = I->learn_backoff;
            I->learn_backoff = std::min(2 * I->learn_backoff, 64);
            I->learn_weak = 0;
        }
    }
    fvp_csr *W = fvp::jacobian_handle(c);
    if (!W) return -1;
    // linear tolerance in the scaled 2-norm; for an ensemble the smallest
    // instance's share, so that every instance meets its own tolerance
    const double tol_abs = I->p.lin_factor * I->p.newton_tol * std::sqrt((double)(I->nseg > 1 ? I->seg_min : N));
    double oldnrm = 0.;
    bool conv = false;
    for (int m = 0; m < I->p.max_newton; ++m) {
        if (m > 0) {
            deriv();
            if (fvp::residual_impl(c, t, I->y.p, I->ypn.p, I->G.p, c->averaging)) return -1;
        }
        // b = -ewt * G ; solve (S J) delta = S (-G), right precond M S^-1
        mul_kernel<<<nb(N), 256, 0, s>>>(N, I->G.p, I->ewt.p, I->b.p);
        axpby_kernel<<<nb(N), 256, 0, s>>>(N, 0., I->b.p, -1., I->b.p);
        FVP_HIP(hipMemsetAsync(I->delta.p, 0, sizeof(double) * N, s));
        double bn = 0.;
        (void)wrms(I, I->G.p, &bn);
        fvp_gmres_params gp{I->p.dim_krylov, I->p.max_restarts, I->p.reorthogonalise, bn > 0 ? tol_abs / bn : 1.};
        fvp_gmres_stats gs{};
        const double *yv = I->y.p, *ypv = I->ypn.p, *Gv = I->G.p, *ew = I->ewt.p;
        double *tmpv = I->tmp.p;
        // ||y||^2 once for the Newton iterate (the FD increment of every J v of this solve)
        double *yn2 = I->scal.p + 4;
        if (dot_async(s, N, yv, yv, I->part.p, yn2)) return -1;
        // S J: the error weights scale the J v rows inside the residual kernels.
        // GMRES applies J to the vector the preconditioner solve has just written
        // (its ||z||^2 partials come with it); any other input is reduced afresh.
        struct LastSolve {
            const double *z = nullptr, *zn2 = nullptr;
            int count = 0;
        } last;
        LastSolve *ls = &last;
        Op A = [=](const double *in, double *out) {
            const bool fresh = in == ls->z && ls->zn2;
            ls->z = nullptr;
            return fvp::jv_impl(c, t, yv, ypv, Gv, cj, in, out, yn2, ew, fresh ? ls->zn2 : nullptr,
                                fresh ? ls->count : 0);
        };
        Op M;
        if (I->p.precon_type != FVP_PRECON_NONE)
            // M S^-1: the block solves divide by the weights as they load
            M = [=](const double *in, double *out) {
                int rc = fvp::precon_solve_div(c, in, ew, tmpv, out, &ls->zn2, &ls->count);
                ls->z = out;
                return rc;
            };
        else
            M = [=](const double *in, double *out) {
                div_kernel<<<nb(N), 256, 0, s>>>(N, in, ew, out);
                return 0;
            };
        int rc = fvp::gmres_device(W, s, N, A, M, I->b.p, I->delta.p, &gp, &gs, true);
        I->st.gmres_iters += gs.iterations;
        if (rc < 0) return -1;
        // a linear solve that stopped short of its tolerance still gives a Newton
        // update when it cut the residual tenfold (inexact Newton: the Newton
        // convergence test judges the result); otherwise a recoverable failure
        // and a smaller step
        if (rc > 0 && !(gs.error < 0.1)) return 2;
        // J v is formed at the current c_j, so the Newton update is taken as
        // solved; IDA scales by 2/(1+c_j ratio) only for a stored (stale)
        // iteration matrix, not for its matrix-free SPGMR solve
        axpby_kernel<<<nb(N), 256, 0, s>>>(N, 1., I->delta.p, 1., I->y.p);
        ++I->st.newton_iters;
        double delnrm = wrms(I, I->delta.p);
        if (m == 0) {
            oldnrm = delnrm;
            if (delnrm <= 1e-4 * I->p.newton_tol || I->ss * delnrm <= I->p.newton_tol) { conv = true; break; }
        } else {
            double rate = std::pow(delnrm / oldnrm, 1.0 / m);
            if (rate > 0.9) break;
            // IDA carries ss into the next step's first iteration; floored here so one
            // fast-contracting step cannot accept a large first update on the next
            I->ss = std::max(rate / (1. - rate), 0.05);
            if (I->ss * delnrm <= I->p.newton_tol) { conv = true; break; }
        }
    }
    if (!conv) return 2;
    deriv();
    if (I->rec_targ) learn_target<<<nb(N), 256, 0, s>>>(N, I->y.p, I->ypred.p, I->ewt.p, I->rec_targ);
    // local error estimate from the predictor-corrector difference
    {
        double c2[2] = {1., -1.};
        const double *x2[2] = {I->y.p, I->ypred.p};
        lincomb(I, 2, c2, x2, I->b.p);
    }
    double fac = (q >= 1 && I->nh >= q + 1) ? h / (tau[0] - tau[q + 1]) : 1.0;
    *err_k = wrms(I, I->b.p) * fac;
    *err_km1 = 1e300;
    if (k >= 2 && I->nh >= k) {
        predict(I, tau, k - 1, I->ypred2.p);
        double c2[2] = {1., -1.};
        const double *x2[2] = {I->y.p, I->ypred2.p};
        lincomb(I, 2, c2, x2, I->b.p);
        *err_km1 = wrms(I, I->b.p) * h / (tau[0] - tau[k]);
    }
    return *err_k <= 1.0 ? 0 : 1;
}

static int accept(fvp_integrator *I, double h) {
    hipStream_t s = I->c->stream;
    // rotate history: the oldest slot receives the new solution
    int oldest = I->slot[NHIST - 1];
    for (int i = NHIST - 1; i > 0; --i) { I->slot[i] = I->slot[i - 1]; I->th[i] = I->th[i - 1]; }
    I->slot[0] = oldest;
    I->th[0] = I->th[1] + h;
    FVP_HIP(hipMemcpyAsync(I->hist[oldest].p, I->y.p, sizeof(double) * I->N, hipMemcpyDeviceToDevice, s));
    FVP_HIP(hipMemcpyAsync(I->yp.p, I->ypn.p, sizeof(double) * I->N, hipMemcpyDeviceToDevice, s));
    I->nh = std::min(I->nh + 1, NHIST);
    ewt_kernel<<<nb(I->N), 256, 0, s>>>(I->N, I->p.rtol, I->p.atol, H(I, 0), I->ewt.p);
    ++I->st.steps;
    ++I->steps_since_setup;
    I->st.t = I->th[0];
    I->st.h_last = h;
    I->st.order_last = I->k;
    return 0;
}

static int one_step(fvp_integrator *I, double t_stop) {
    fvp_ctx *c = I->c;
    // freeze the upwind weights with the state at t_n (preprocess_timestep)
    if (fvp_begin_step(c, I->th[0], H(I, 0), I->yp.p)) return -1;
    int fails = 0;
    for (;;) {
        double h = I->h;
        if (I->p.h_max > 0) h = std::min(h, I->p.h_max);
        bool clamp = false;
        if (I->th[0] + h >= t_stop) { h = t_stop - I->th[0]; clamp = true; }
        if (!(h > 1e-14 * std::max(1.0, std::fabs(I->th[0])))) { set_error("step size underflow"); return 3; }
        int k = std::min(I->k, std::max(1, I->nh));
        double ek = 0., ekm1 = 0.;
        int rc = attempt(I, h, k, &ek, &ekm1);
        if (rc < 0) return -1;
        if (rc == 0) {
            if (accept(I, h)) return -1;
            I->learn_block = false;
            // order and step selection
            int knew = k;
            double enew = ek;
            if (k >= 2 && ekm1 <= ek) { knew = k - 1; enew = ekm1; }
            else if (k < I->p.max_order && I->steps_at_k >= k + 1 && I->nh >= k + 2 && ek < 0.5) knew = k + 1;
            I->steps_at_k = knew == I->k ? I->steps_at_k + 1 : 0;
            I->k = knew;
            double ratio = std::pow(2. * enew + 1e-4, -1. / (knew + 1));
            double hn = h;
            if (ratio >= 2.) hn = 2. * h;
            else if (ratio <= 1.) hn = h * std::max(0.5, std::min(0.9, ratio));
            // a step cut short by t_stop does not shrink the next one
            I->h = clamp ? std::max(I->h, hn) : hn;
            return 0;
        }
        ++fails;
        if (rc == 1) {
            ++I->st.error_fails;
            double r = 0.9 * std::pow(ek, -1. / (k + 1));
            I->h = h * std::max(0.25, std::min(0.9, r));
            if (fails >= 3) { I->k = 1; I->steps_at_k = 0; }
        } else {
            ++I->st.newton_fails;
            // fallback: a step whose Newton iteration failed from the network's
            // start is retried from the predictor
            if (I->learn_took) I->learn_block = true;
            I->h = h * 0.25;
            I->need_setup = true;
        }
        if (fails > 20) { set_error("too many step failures"); return 3; }
    }
}

}  // namespace fvp

using namespace fvp;

extern "C" int fvp_integrator_create(fvp_ctx *c, const fvp_integrator_params *p, double t0, const double *psi0,
                                     fvp_integrator **out) {
    if (!c || !p || !psi0 || !out) { set_error("null argument"); return -1; }
    if (p->max_order < 1 || p->max_order > KMAX) { set_error("max_order must be 1..5"); return -1; }
    if (!(p->rtol >= 0) || !(p->atol > 0)) { set_error("tolerances must be positive"); return -1; }
    if (p->dim_krylov < 1 || p->dim_krylov > 63) { set_error("dim_krylov must be 1..63"); return -1; }
    if (p->reorthogonalise < 0 || p->reorthogonalise > 2) { set_error("reorthogonalise must be 0, 1 or 2"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    fvp_integrator *I = new fvp_integrator();
    I->c = c;
    I->p = *p;
    if (I->p.max_newton <= 0) I->p.max_newton = 4;
    if (I->p.newton_tol <= 0) I->p.newton_tol = 0.33;
    if (I->p.lin_factor <= 0) I->p.lin_factor = 0.05;
    if (I->p.max_steps <= 0) I->p.max_steps = 100000;
    if (I->p.block_rows > 0 && fvp_set_block_rows(c, I->p.block_rows)) { delete I; return -1; }
    if (I->p.overlap >= 0 && fvp_set_overlap(c, I->p.overlap)) { delete I; return -1; }
    I->N = c->N;
    const long N = c->N;
    for (int i = 0; i < NHIST; ++i) {
        if (I->hist[i].alloc(N)) { fvp_integrator_destroy(I); return -1; }
        I->slot[i] = i;
        I->th[i] = t0;
    }
    if (I->yp.alloc(N) || I->ewt.alloc(N) || I->y.alloc(N) || I->ypn.alloc(N) || I->ypred.alloc(N) ||
        I->ypred2.alloc(N) || I->G.alloc(N) || I->b.alloc(N) || I->delta.alloc(N) || I->beta.alloc(N) ||
        I->tmp.alloc(N) || I->part.alloc(512) || I->scal.alloc(8)) {
        fvp_integrator_destroy(I);
        return -1;
    }
    if (hipHostMalloc((void **)&I->hscal, sizeof(double) * 8, hipHostMallocDefault) != hipSuccess) {
        I->hscal = nullptr;
        fvp_integrator_destroy(I);
        set_error("pinned host allocation failed");
        return -1;
    }
    if (hipEventCreateWithFlags(&I->hev, hipEventDisableTiming) != hipSuccess) {
        I->hev = nullptr;
        fvp_integrator_destroy(I);
        set_error("event creation failed");
        return -1;
    }
    hipStream_t s = c->stream;
    FVP_HIP(hipMemcpyAsync(I->hist[0].p, psi0, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    I->nh = 1;
    // consistent initial derivative: residual with psi' = 0 (averaged weights), psi' = F / ahh
    FVP_HIP(hipMemsetAsync(I->yp.p, 0, sizeof(double) * N, s));
    if (residual_impl(c, t0, I->hist[0].p, I->yp.p, I->G.p, c->averaging)) return -1;
    ic_deriv<<<nb(N), 256, 0, s>>>(N, I->G.p, c->cap.p, I->yp.p);
    if (c->ndir) zero_rows<<<(c->ndir + 255) / 256, 256, 0, s>>>(c->ndir, c->dnode.p, I->yp.p);
    ewt_kernel<<<nb(N), 256, 0, s>>>(N, I->p.rtol, I->p.atol, I->hist[0].p, I->ewt.p);
    FVP_HIP(hipGetLastError());
    I->st.t = t0;
    I->h = p->h0;
    I->nres0 = c->nres;
    *out = I;
    return 0;
}

extern "C" int fvp_integrator_destroy(fvp_integrator *I) {
    if (!I) return 0;
    for (int i = 0; i < NHIST; ++i) I->hist[i].release();
    I->yp.release(); I->ewt.release(); I->y.release(); I->ypn.release(); I->ypred.release(); I->ypred2.release();
    I->G.release(); I->b.release(); I->delta.release(); I->beta.release(); I->tmp.release(); I->part.release();
    I->scal.release();
    I->seg.release(); I->seg_sum.release();
    I->stat.release(); I->feat.release(); I->ylearn.release(); I->y2.release(); I->ypn2.release();
    I->G2.release();
    if (I->hscal) (void)hipHostFree(I->hscal);
    if (I->hev) (void)hipEventDestroy(I->hev);
    delete I;
    return 0;
}

static int init_h(fvp_integrator *I, double t_out) {
    if (I->h > 0) return 0;
    double span = std::fabs(t_out - I->th[0]);
    double ypn = wrms(I, I->yp.p);
    double h = 1e-3 * span;
    if (ypn * h > 0.5) h = 0.5 / ypn;
    I->h = std::max(h, 1e-10 * std::max(1.0, span));
    return 0;
}

extern "C" int fvp_integrator_step(fvp_integrator *I, double t_stop) {
    if (!I) { set_error("null integrator"); return -1; }
    FVP_HIP(hipSetDevice(I->c->device));
    if (!(t_stop > I->th[0])) { set_error("t_stop must lie ahead of the current time"); return -1; }
    init_h(I, t_stop);
    return one_step(I, t_stop);
}

extern "C" int fvp_integrator_advance(fvp_integrator *I, double t_out, double *psi_out) {
    if (!I) { set_error("null integrator"); return -1; }
    FVP_HIP(hipSetDevice(I->c->device));
    init_h(I, t_out);
    int n = 0;
    while (I->th[0] < t_out) {
        int rc = one_step(I, t_out);
        if (rc) return rc;
        if (++n > I->p.max_steps) { set_error("max_steps reached"); return 3; }
    }
    if (psi_out)
        FVP_HIP(hipMemcpyAsync(psi_out, H(I, 0), sizeof(double) * I->N, hipMemcpyDeviceToDevice, I->c->stream));
    return 0;
}

extern "C" int fvp_integrator_set_segments(fvp_integrator *I, int nseg, const int *offsets) {
    if (!I || (nseg > 0 && !offsets)) { set_error("null argument"); return -1; }
    if (nseg < 0) { set_error("nseg must be >= 0"); return -1; }
    if (nseg == 0) { I->nseg = 0; return 0; }
    if (offsets[0] != 0 || offsets[nseg] != I->N) { set_error("segments must start at 0 and end at n_nodes"); return -1; }
    int mn = I->N;
    for (int g = 0; g < nseg; ++g) {
        if (offsets[g + 1] <= offsets[g]) { set_error("segment offsets must increase"); return -1; }
        mn = std::min(mn, offsets[g + 1] - offsets[g]);
    }
    if (nseg == 1) { I->nseg = 0; return 0; }  // one instance: the plain norms
    FVP_HIP(hipSetDevice(I->c->device));
    if (I->seg.upload(offsets, nseg + 1) || I->seg_sum.alloc(nseg)) return -1;
    I->nseg = nseg;
    I->seg_min = mn;
    return 0;
}

// static feature columns (device N x 16) into the integrator's copy; null = zeros
static int set_static(fvp_integrator *I, const float *static_features) {
    const long N = I->N;
    if (!I->stat.p && I->stat.alloc(N * 16)) return -1;
    hipStream_t s = I->c->stream;
    if (static_features)
        FVP_HIP(hipMemcpyAsync(I->stat.p, static_features, sizeof(float) * N * 16, hipMemcpyDeviceToDevice, s));
    else
        FVP_HIP(hipMemsetAsync(I->stat.p, 0, sizeof(float) * N * 16, s));
    return 0;
}

extern "C" int fvp_integrator_set_learned(fvp_integrator *I, fvp_mpnn *m, const float *static_features, double trust,
                                          double band) {
    if (!I) { set_error("null integrator"); return -1; }
    if (m && (!(trust >= 0.) || !(band >= 0.))) { set_error("trust and band must be >= 0"); return -1; }
    FVP_HIP(hipSetDevice(I->c->device));
    if (!m) {
        I->mp = nullptr;
        I->feat.release(); I->ylearn.release(); I->y2.release(); I->ypn2.release(); I->G2.release();
        return 0;
    }
    if (fvp::mpnn_nodes(m) != I->N) { set_error("network node count differs from the problem's"); return -1; }
    const long N = I->N;
    if (set_static(I, static_features) || (!I->feat.p && I->feat.alloc(N * 16)) || I->ylearn.alloc(N) ||
        I->y2.alloc(N) || I->ypn2.alloc(N) || I->G2.alloc(N))
        return -1;
    if (fvp_mpnn_set_stream(m, (void *)I->c->stream)) return -1;
    I->mp = m;
    I->learn_trust = trust;
    I->learn_band = band;
    I->learn_weak = I->learn_skip = 0;
    I->learn_backoff = 4;
    return 0;
}

extern "C" int fvp_integrator_set_recording(fvp_integrator *I, const float *static_features, float *features,
                                            float *targets) {
    if (!I) { set_error("null integrator"); return -1; }
    if ((features == nullptr) != (targets == nullptr)) { set_error("give both recording buffers or neither"); return -1; }
    FVP_HIP(hipSetDevice(I->c->device));
    if (features && set_static(I, static_features)) return -1;
    I->rec_feat = features;
    I->rec_targ = targets;
    return 0;
}

extern "C" int fvp_integrator_state(fvp_integrator *I, double *t, const double **psi, const double **psip) {
    if (!I) { set_error("null integrator"); return -1; }
    if (t) *t = I->th[0];
    if (psi) *psi = H(I, 0);
    if (psip) *psip = I->yp.p;
    return 0;
}

extern "C" int fvp_integrator_get_stats(const fvp_integrator *I, fvp_integrator_stats *st) {
    if (!I || !st) { set_error("null argument"); return -1; }
    *st = I->st;
    st->residuals = I->c->nres - I->nres0;
    return 0;
}
