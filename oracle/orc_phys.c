/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle: plain-C restatement of the
 * reference's variably-saturated CVFE residual (VarSatPhysics).  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it; the product library never links or calls it.
 *
 * Followed reference files (paths relative to the reference root):
 *   examples/fvmporMinlin/fvmpor.h:280-344    saturation (host twin of the GPU kernel)
 *   examples/fvmporMinlin/fvmpor.h:239-278    density / porosity
 *   examples/fvmporMinlin/fvmpor.h:346-644    initialise_vectors (zones, SCV weights, Dirichlet)
 *   examples/fvmporMinlin/fvmpor.h:653-679    process_faces_lim
 *   examples/fvmporMinlin/fvmpor.h:681-759    process_fluxes
 *   examples/fvmporMinlin/fvmpor.h:761-797    process_spatial_weights (upwind)
 *   examples/fvmporMinlin/fvmpor.h:874-945    process_volumes_psk / derivative coefficients
 *   examples/fvmporMinlin/fvmpor_ODE_impl.cpp:62-142  preprocess / residual / timestep
 *   examples/fvmporMinlin/definitions.h:105-200       boundary-condition types
 *
 * Extension: a specific-storage term.  The capacity coefficient is
 *   ahh = rho0 * (phi * dSw + Ss * Sw)       (CV-averaged phi, dSw, Ss, Sw)
 * With Ss = 0 every value equals the reference's ahh = rho0 * phi * dSw
 * (fvmpor.h:943-944) bit for bit.
 *
 * Per-element conductivity: K on a CV face is -(K_zone * kmul_element), the
 * reference's per-zone K_faces_ (fvmpor.h:632-643) with kmul = 1.
 *
 * Parity pins (tests/test_oracle_physics.py): the van Genuchten closed forms
 * and special points, a finite-difference check of dSw/dh, hydrostatic
 * equilibrium (F = 0 to rounding), global mass conservation and the linear
 * saturated field.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { ZP_KXX, ZP_KYY, ZP_KZZ, ZP_PHI, ZP_ALPHA, ZP_N, ZP_SR, ZP_SS, ZP_COUNT };
enum { BC_VALUE, BC_B, BC_ETA, BC_QX, BC_QY, BC_QZ, BC_COUNT };

typedef struct {
    OrcMesh *m;
    int nz;
    double *zp;     /* nz x ZP_COUNT */
    int *ezone;     /* T */
    double *ekmul;  /* T */
    int ntags;
    int *bc_type;   /* ntags (tag index = boundary tag value) */
    double *bc;     /* ntags x BC_COUNT */
    double rho0;
    /* zone membership: zone z holds node i with weight w (fvmpor.h:466-498) */
    int *zn_ptr, *zn_node;
    double *zn_w;
    int *zn_slot;   /* node -> position in its zone list, per zone: nz x N, -1 if absent */
    /* Dirichlet rows */
    int *dir_tag;   /* N: last Dirichlet tag of the node, 0 otherwise */
    int ndir;
    int *dir_node;
    double *dir_val;
    /* state */
    double *wf, *wb;                /* NE edge weights */
    double *kr_zn;                  /* per zone list entry */
    double *gx, *gy, *gz, *krw;     /* NFI */
    double *qdotn, *M;              /* NFI + NFB */
    double *phi_v, *Sw_v, *dSw_v, *theta_v, *Ss_v, *ahh; /* N */
    double *edge_flux;              /* NE */
    double *hp_last;
} OrcProblem;

/* minlin host twin of lin::gpu::saturation (fvmpor.h:302-343) */
static void vg_twin(double h, const double *zp, double *Sw, double *dSw, double *krw) {
    double alphaVG = zp[ZP_ALPHA], nVG = zp[ZP_N], mVG = 1. - 1. / zp[ZP_N], S_r = zp[ZP_SR];
    if (h >= 0.) { *dSw = 0.; *Sw = 1.; *krw = 1.; return; }
    double d = -alphaVG * h;
    d = pow(d, nVG);          /* a */
    double s = d + 1.;        /* b */
    double k = -1.;
    k /= s;                   /* -1/b */
    d /= s;                   /* a/b */
    s = pow(s, -mVG);         /* Se */
    d *= s;
    d /= h;
    d *= -(1 - S_r) * (nVG - 1);
    k += 1.;
    k = pow(k, mVG);
    k -= 1.;
    k = pow(k, 2);
    k *= sqrt(s);
    s *= (1 - S_r);
    s += S_r;
    *Sw = s; *dSw = d; *krw = k;
}

void orc_vg(double h, const double *zp, double *out3) { vg_twin(h, zp, out3, out3 + 1, out3 + 2); }

OrcProblem *orc_problem_create(OrcMesh *m, int nz, const double *zp, const int *ezone,
                               const double *ekmul, int ntags, const int *bc_type, const double *bc,
                               const int *ntag_ptr, const int *ntag, double rho0) {
    OrcProblem *p = (OrcProblem *)calloc(1, sizeof(OrcProblem));
    int N = m->N, T = m->T, npe = m->npe, NE = m->NE, NFI = m->NFI, NF = m->NFI + m->NFB;
    p->m = m; p->nz = nz; p->rho0 = rho0; p->ntags = ntags;
    p->zp = (double *)malloc(sizeof(double) * nz * ZP_COUNT);
    memcpy(p->zp, zp, sizeof(double) * nz * ZP_COUNT);
    p->ezone = (int *)malloc(sizeof(int) * T);
    memcpy(p->ezone, ezone, sizeof(int) * T);
    p->ekmul = (double *)malloc(sizeof(double) * T);
    memcpy(p->ekmul, ekmul, sizeof(double) * T);
    p->bc_type = (int *)malloc(sizeof(int) * (ntags > 0 ? ntags : 1));
    p->bc = (double *)malloc(sizeof(double) * BC_COUNT * (ntags > 0 ? ntags : 1));
    if (ntags) {
        memcpy(p->bc_type, bc_type, sizeof(int) * ntags);
        memcpy(p->bc, bc, sizeof(double) * BC_COUNT * ntags);
    }

    /* zone lists with SCV-volume weights, zones in ascending index, nodes ascending */
    double *w = (double *)calloc((size_t)nz * N, sizeof(double));
    int *cnt = (int *)calloc((size_t)nz * N, sizeof(int));
    for (int e = 0; e < T; ++e)
        for (int q = 0; q < npe; ++q) {
            int i = m->conn[e * npe + q], z = ezone[e];
            cnt[(size_t)z * N + i]++;
        }
    /* weights accumulate over the node's SCVs in element order (fvmpor.h:478-482) */
    for (int e = 0; e < T; ++e)
        for (int q = 0; q < npe; ++q) {
            int i = m->conn[e * npe + q], z = ezone[e];
            w[(size_t)z * N + i] += m->scv_vol[e * npe + q] / m->V[i];
        }
    p->zn_ptr = (int *)calloc(nz + 1, sizeof(int));
    for (int z = 0; z < nz; ++z) {
        int c = 0;
        for (int i = 0; i < N; ++i) c += cnt[(size_t)z * N + i] > 0;
        p->zn_ptr[z + 1] = p->zn_ptr[z] + c;
    }
    int tot = p->zn_ptr[nz];
    p->zn_node = (int *)malloc(sizeof(int) * (tot ? tot : 1));
    p->zn_w = (double *)malloc(sizeof(double) * (tot ? tot : 1));
    p->zn_slot = (int *)malloc(sizeof(int) * (size_t)nz * N);
    p->kr_zn = (double *)malloc(sizeof(double) * (tot ? tot : 1));
    for (int z = 0, k = 0; z < nz; ++z)
        for (int i = 0; i < N; ++i) {
            if (cnt[(size_t)z * N + i]) {
                p->zn_node[k] = i; p->zn_w[k] = w[(size_t)z * N + i];
                p->zn_slot[(size_t)z * N + i] = k++;
            } else p->zn_slot[(size_t)z * N + i] = -1;
        }
    free(w); free(cnt);

    /* Dirichlet tagging (fvmpor.h:404-451) */
    p->dir_tag = (int *)calloc(N, sizeof(int));
    int nd = 0;
    for (int i = 0; i < N; ++i) {
        for (int k = ntag_ptr[i]; k < ntag_ptr[i + 1]; ++k) {
            int tag = ntag[k];
            if (tag > 0 && tag < ntags && (bc_type[tag] == 1 || bc_type[tag] == 4)) {
                if (!p->dir_tag[i]) ++nd;
                p->dir_tag[i] = tag;
            }
        }
    }
    p->ndir = nd;
    p->dir_node = (int *)malloc(sizeof(int) * (nd ? nd : 1));
    p->dir_val = (double *)malloc(sizeof(double) * (nd ? nd : 1));
    for (int i = 0, k = 0; i < N; ++i)
        if (p->dir_tag[i]) {
            int tag = p->dir_tag[i];
            const double *b = p->bc + BC_COUNT * tag;
            double el = m->dim == 2 ? m->xyz[3 * i + 1] : m->xyz[3 * i + 2];
            p->dir_node[k] = i;
            p->dir_val[k] = bc_type[tag] == 1 ? b[BC_VALUE] : (1. + b[BC_ETA]) * (b[BC_B] - el);
            ++k;
        }

    p->wf = (double *)malloc(sizeof(double) * NE);
    p->wb = (double *)malloc(sizeof(double) * NE);
    for (int i = 0; i < NE; ++i) { p->wf[i] = 0.5; p->wb[i] = 0.5; }
    p->gx = (double *)calloc(NFI, sizeof(double));
    p->gy = (double *)calloc(NFI, sizeof(double));
    p->gz = (double *)calloc(NFI, sizeof(double));
    p->krw = (double *)calloc(NFI, sizeof(double));
    p->qdotn = (double *)calloc(NF, sizeof(double));
    p->M = (double *)calloc(NF, sizeof(double));
    p->phi_v = (double *)calloc(N, sizeof(double));
    p->Sw_v = (double *)calloc(N, sizeof(double));
    p->dSw_v = (double *)calloc(N, sizeof(double));
    p->theta_v = (double *)calloc(N, sizeof(double));
    p->Ss_v = (double *)calloc(N, sizeof(double));
    p->ahh = (double *)calloc(N, sizeof(double));
    p->edge_flux = (double *)calloc(NE, sizeof(double));
    return p;
}

void orc_problem_free(OrcProblem *p) {
    if (!p) return;
    free(p->zp); free(p->ezone); free(p->ekmul); free(p->bc_type); free(p->bc);
    free(p->zn_ptr); free(p->zn_node); free(p->zn_w); free(p->zn_slot); free(p->kr_zn);
    free(p->dir_tag); free(p->dir_node); free(p->dir_val); free(p->wf); free(p->wb);
    free(p->gx); free(p->gy); free(p->gz); free(p->krw); free(p->qdotn); free(p->M);
    free(p->phi_v); free(p->Sw_v); free(p->dSw_v); free(p->theta_v); free(p->Ss_v); free(p->ahh);
    free(p->edge_flux);
    free(p);
}

/* y = A x over a CSR matrix, plain loop (interpolation_legacy.h:30-36) */
static void csr_mv(int nrows, const int *ia, const int *ja, const double *a, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nrows; ++i) {
        double s = 0.;
        for (int k = ia[i]; k < ia[i + 1]; ++k) s += a[k] * x[ja[k]];
        y[i] = s;
    }
}

/* preprocess_evaluation (fvmpor_ODE_impl.cpp:62-97) */
static void preprocess(OrcProblem *p, double t, const double *h) {
    OrcMesh *m = p->m;
    int N = m->N, NFI = m->NFI, NF = m->NFI + m->NFB, nz = p->nz;
    (void)t;
    csr_mv(NFI, m->sh_ia, m->sh_ja, m->sh_dx, h, p->gx);
    csr_mv(NFI, m->sh_ia, m->sh_ja, m->sh_dy, h, p->gy);
    if (m->dim == 3) csr_mv(NFI, m->sh_ia, m->sh_ja, m->sh_dz, h, p->gz);

    /* process_volumes_psk (fvmpor.h:874-928) */
    memset(p->phi_v, 0, sizeof(double) * N);
    memset(p->Sw_v, 0, sizeof(double) * N);
    memset(p->dSw_v, 0, sizeof(double) * N);
    memset(p->theta_v, 0, sizeof(double) * N);
    memset(p->Ss_v, 0, sizeof(double) * N);
    for (int z = 0; z < nz; ++z) {
        const double *zp = p->zp + z * ZP_COUNT;
        double phi = zp[ZP_PHI]; /* porosity() with alpha == 0 (fvmpor.h:269-272) */
        /* a node appears once per zone: the zone's nodes update disjoint entries */
#pragma omp parallel for schedule(static)
        for (int k = p->zn_ptr[z]; k < p->zn_ptr[z + 1]; ++k) {
            int i = p->zn_node[k];
            double wgt = p->zn_w[k];
            double Sw, dSw, kr;
            vg_twin(h[i], zp, &Sw, &dSw, &kr);
            double theta = Sw * phi;
            p->kr_zn[k] = kr;
            p->phi_v[i] += phi * wgt;
            p->Sw_v[i] += Sw * wgt;
            p->dSw_v[i] += dSw * wgt;
            p->theta_v[i] += theta * wgt;
            p->Ss_v[i] += zp[ZP_SS] * wgt;
        }
    }
    /* k_r at faces from the face element's zone, front then back (fvmpor.h:919-924) */
#pragma omp parallel for schedule(static)
    for (int f = 0; f < NFI; ++f) {
        const CvFace *c = &m->cf[f];
        int z = p->ezone[c->elem];
        double kf = p->kr_zn[p->zn_slot[(size_t)z * N + c->front]];
        double kb = p->kr_zn[p->zn_slot[(size_t)z * N + c->back]];
        double v = kf * p->wf[c->edge];
        v += kb * p->wb[c->edge];
        p->krw[f] = v;
    }
    /* process_derivative_coefficients (fvmpor.h:930-945) + specific storage */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) p->ahh[i] = (p->phi_v[i] * p->dSw_v[i] + p->Ss_v[i] * p->Sw_v[i]) * p->rho0;

    /* process_fluxes (fvmpor.h:681-759) */
#pragma omp parallel for schedule(static)
    for (int f = 0; f < NFI; ++f) {
        const CvFace *c = &m->cf[f];
        int z = p->ezone[c->elem];
        const double *zp = p->zp + z * ZP_COUNT;
        double km = p->ekmul[c->elem];
        double Kx = -(zp[ZP_KXX] * km), Ky = -(zp[ZP_KYY] * km), Kz = -(zp[ZP_KZZ] * km);
        double qx = p->gx[f] * Kx;
        double qy = p->gy[f], qz = 0.;
        if (m->dim == 2) qy += 1.;
        else { qz = p->gz[f]; qz += 1.; qz *= Kz; }
        qy *= Ky;
        double nx = c->nx * c->area, ny = c->ny * c->area, nzv = c->nz * c->area;
        double q = nx * qx;
        q += ny * qy;
        if (m->dim == 3) q += nzv * qz;
        q *= p->krw[f];
        p->qdotn[f] = q;
        /* rho_lim from the edge weights with rho = rho0 (beta = 0; fvmpor.h:664-678) */
        double rho_edge = p->rho0 * p->wb[c->edge] + p->rho0 * p->wf[c->edge];
        p->M[f] = rho_edge * q;
    }
    for (int f = NFI; f < NF; ++f) {
        const CvFace *c = &m->cf[f];
        int tag = c->tag;
        int type = (tag >= 0 && tag < p->ntags) ? p->bc_type[tag] : -1;
        const double *b = p->bc + BC_COUNT * (tag >= 0 && tag < p->ntags ? tag : 0);
        double q = 0.;
        switch (type) {
        case 3: q = b[BC_VALUE] * c->area; break;
        case 6: { /* flux(t, normal()) * area, normal() = n * area (definitions.h:166-176) */
            double nx = c->nx * c->area, ny = c->ny * c->area, nzv = c->nz * c->area;
            q = (nx * b[BC_QX] + ny * b[BC_QY] + nzv * b[BC_QZ]) * c->area;
            break;
        }
        case 7: q = b[BC_VALUE] * c->area; break;
        case 8: q = 0. * c->area; break;
        default: q = 0.; break;
        }
        p->qdotn[f] = q;
        p->M[f] = p->rho0 * q;
    }
}

/* residual_evaluation (fvmpor_ODE_impl.cpp:99-115) */
void orc_residual(OrcProblem *p, double t, const double *h, const double *hp, double *res) {
    OrcMesh *m = p->m;
    int N = m->N;
    preprocess(p, t, h);
    csr_mv(N, m->cv_ia, m->cv_ja, m->cv_w, p->M, res);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) res[i] -= hp[i] * p->ahh[i];
    for (int k = 0; k < p->ndir; ++k) {
        int i = p->dir_node[k];
        res[i] = p->dir_val[k];
        res[i] -= h[i];
    }
}

/* preprocess_timestep + process_spatial_weights, upwind case (fvmpor_ODE_impl.cpp:118-142;
 * fvmpor.h:761-797) */
void orc_begin_step(OrcProblem *p, double t, const double *h, const double *hp) {
    OrcMesh *m = p->m;
    (void)hp;
    for (int i = 0; i < m->NE; ++i) { p->wf[i] = 0.5; p->wb[i] = 0.5; }
    preprocess(p, t, h);
    csr_mv(m->NE, m->fl_ia, m->fl_ja, m->fl_w, p->qdotn, p->edge_flux);
    for (int i = 0; i < m->NE; ++i) {
        if (p->edge_flux[i] < 0.) { p->wf[i] = 1.; p->wb[i] = 0.; }
        else { p->wf[i] = 0.; p->wb[i] = 1.; }
    }
}

void orc_set_weights(OrcProblem *p, const double *wf) {
    for (int i = 0; i < p->m->NE; ++i) { p->wf[i] = wf[i]; p->wb[i] = 1. - wf[i]; }
}
void orc_get_weights(const OrcProblem *p, double *wf, double *edge_flux) {
    memcpy(wf, p->wf, sizeof(double) * p->m->NE);
    if (edge_flux) memcpy(edge_flux, p->edge_flux, sizeof(double) * p->m->NE);
}

/* internals after the last evaluation: ahh, Sw, dSw (N each) */
void orc_get_node_fields(const OrcProblem *p, double *ahh, double *Sw, double *dSw) {
    int N = p->m->N;
    if (ahh) memcpy(ahh, p->ahh, sizeof(double) * N);
    if (Sw) memcpy(Sw, p->Sw_v, sizeof(double) * N);
    if (dSw) memcpy(dSw, p->dSw_v, sizeof(double) * N);
}
/* qdotn and M over all CV faces after the last evaluation */
void orc_get_face_fields(const OrcProblem *p, double *qdotn, double *M) {
    int NF = p->m->NFI + p->m->NFB;
    if (qdotn) memcpy(qdotn, p->qdotn, sizeof(double) * NF);
    if (M) memcpy(M, p->M, sizeof(double) * NF);
}

/* flux scale of the parity definition: max_i (1/V_i) * sum_f |M_f| over the CV faces of i */
double orc_flux_scale(const OrcProblem *p) {
    const OrcMesh *m = p->m;
    double s = 0.;
    for (int i = 0; i < m->N; ++i) {
        double a = 0.;
        for (int k = m->cv_ia[i]; k < m->cv_ia[i + 1]; ++k) a += fabs(p->M[m->cv_ja[k]]);
        a /= m->V[i];
        if (a > s) s = a;
    }
    return s;
}

int orc_num_dirichlet(const OrcProblem *p) { return p->ndir; }
void orc_dirichlet(const OrcProblem *p, int *nodes, double *vals) {
    memcpy(nodes, p->dir_node, sizeof(int) * p->ndir);
    memcpy(vals, p->dir_val, sizeof(double) * p->ndir);
}

OrcMesh *orc_problem_mesh(OrcProblem *p) { return p->m; }
