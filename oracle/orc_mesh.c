/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle: a plain-C restatement of the
 * reference's CVFE mesh dual (control volumes, CV faces, SCV volumes, shape
 * function gradients and the interpolation matrices built on them).  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it; the product library never links or calls it.
 *
 * Followed reference files (paths relative to the reference root):
 *   src/fvm/mesh.cpp                      mesh construction
 *   include/fvm/impl/mesh/cvface_shape.h  CV face area / centroid / normal
 *   include/fvm/impl/mesh/cvface_inline.h CV face orientation and ordering
 *   include/fvm/impl/mesh/edge_inline.h   edge ordering
 *   examples/fvmporMinlin/shape.cpp       shape-function gradients
 *   examples/fvmporMinlin/fvmpor.h        initialise_shape_functions
 *
 * Deliberate differences, none of which changes a value beyond rounding:
 *   - no RCM renumbering (mesh.cpp:118-209): nodes keep the caller's order.
 *     Edge front/back roles do not depend on numbering (the first element
 *     that inserts an edge fixes them, mesh.cpp:385-395 + std::set::insert),
 *     and edges are still sorted by (min, max) node id as mesh.cpp:163 does.
 *   - the signed pyramid height of mesh.cpp:1519-1535 is evaluated as
 *     dot(c, apex-P1) / dot(c, n) with c = (P2-P1)x(P3-P1), the same
 *     determinant written with cross products.
 * Parity pins for this file live in tests/test_oracle_mesh.py: SCV volumes
 * against |T|/(d+1), sum of CV volumes against the domain volume, closed CV
 * surfaces, back->front normals, exact gradients of linear fields and the
 * reference's own Jacobian pattern for the cassion mesh (meshing/Jpattern.mtx).
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- small vector helpers ------------------------------------------- */
typedef struct { double x, y, z; } P3;
static P3 p3(double x, double y, double z) { P3 p = {x, y, z}; return p; }
static P3 add(P3 a, P3 b) { return p3(a.x + b.x, a.y + b.y, a.z + b.z); }
static P3 sub(P3 a, P3 b) { return p3(a.x - b.x, a.y - b.y, a.z - b.z); }
static P3 scl(P3 a, double s) { return p3(a.x * s, a.y * s, a.z * s); }
static P3 dvd(P3 a, double s) { return p3(a.x / s, a.y / s, a.z / s); }
static double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static P3 cross(P3 a, P3 b) {
    return p3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static double nrm(P3 a) { return sqrt(dot(a, a)); }

static P3 node_pt(const OrcMesh *m, int i) {
    return p3(m->xyz[3 * i], m->xyz[3 * i + 1], m->xyz[3 * i + 2]);
}
P3 orc_face_centroid(const OrcMesh *m, int src);

/* element-local edge and face tables (mesh.cpp:409-500) */
static const int TRI_EDGE[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int TET_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static const int TRI_FACE_NODES[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int TET_FACE_NODES[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
static const int TET_FACE_EDGES[4][3] = {{0, 1, 2}, {0, 4, 3}, {2, 5, 3}, {1, 5, 4}};

/* ---- CV face shape (cvface_shape.h) ----------------------------------- */
typedef struct { P3 p[4]; int np; } Shape;

static double shape_area(const Shape *s) {
    if (s->np == 2) return nrm(sub(s->p[1], s->p[0]));
    return (nrm(cross(sub(s->p[1], s->p[0]), sub(s->p[3], s->p[0]))) +
            nrm(cross(sub(s->p[1], s->p[2]), sub(s->p[3], s->p[2])))) / 2.0;
}
static P3 shape_centroid(const Shape *s) {
    if (s->np == 2) return dvd(add(s->p[0], s->p[1]), 2.0);
    P3 c1 = dvd(add(add(s->p[0], s->p[1]), s->p[3]), 3.0);
    P3 c2 = dvd(add(add(s->p[1], s->p[2]), s->p[3]), 3.0);
    double a1 = nrm(cross(sub(s->p[1], s->p[0]), sub(s->p[3], s->p[0]))) / 2.0;
    double a2 = nrm(cross(sub(s->p[1], s->p[2]), sub(s->p[3], s->p[2]))) / 2.0;
    return dvd(add(scl(c1, a1), scl(c2, a2)), a1 + a2);
}
static P3 shape_normal(const Shape *s) {
    P3 r;
    if (s->np == 2) r = cross(sub(s->p[1], s->p[0]), p3(0, 0, 1));
    else r = cross(sub(s->p[1], s->p[0]), sub(s->p[3], s->p[0]));
    return dvd(r, nrm(r));
}
static void shape_reverse(Shape *s) {
    P3 t;
    if (s->np == 2) { t = s->p[0]; s->p[0] = s->p[1]; s->p[1] = t; }
    else { t = s->p[1]; s->p[1] = s->p[3]; s->p[3] = t; }
}

/* mesh.cpp:1488-1498 */
static double triangle_volume(P3 p1, P3 p2, P3 p3v) {
    P3 v1 = sub(p1, p2), v2 = sub(p1, p3v);
    double a = dot(v1, v1), b = dot(v2, v2), c = dot(v1, v2);
    return sqrt(a * b - c * c) / 2.0;
}
/* mesh.cpp:1500-1539 */
static double pyramid_volume(const Shape *f, P3 apex) {
    P3 P1 = f->p[0], P2 = f->p[1], P3v = f->p[3];
    P3 n = shape_normal(f);
    P3 c = cross(sub(P2, P1), sub(P3v, P1));
    double height = dot(c, sub(apex, P1)) / dot(c, n);
    return fabs(shape_area(f) * height / 3.0);
}

/* ---- sorting helpers ------------------------------------------------- */
typedef struct { int a, b, c, d; } I4;
static int cmp_i4(const void *x, const void *y) {
    const I4 *p = (const I4 *)x, *q = (const I4 *)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    if (p->c != q->c) return p->c < q->c ? -1 : 1;
    if (p->d != q->d) return p->d < q->d ? -1 : 1;
    return 0;
}
static int cmp_int(const void *x, const void *y) {
    int a = *(const int *)x, b = *(const int *)y;
    return a < b ? -1 : (a > b);
}

/* ---- construction ---------------------------------------------------- */
static int nfe_of(int dim) { return dim == 2 ? 3 : 4; }
static int ned_of(int dim) { return dim == 2 ? 3 : 6; }

OrcMesh *orc_mesh_create(int dim, int N, const double *xyz, int T, const int *conn,
                         const int *eface_tag, int *status) {
    *status = 0;
    if (dim != 2 && dim != 3) { *status = -1; return NULL; }
    OrcMesh *m = (OrcMesh *)calloc(1, sizeof(OrcMesh));
    int npe = dim + 1, ned = ned_of(dim), nfe = nfe_of(dim);
    m->dim = dim; m->N = N; m->T = T; m->npe = npe;
    m->xyz = (double *)malloc(sizeof(double) * 3 * N);
    memcpy(m->xyz, xyz, sizeof(double) * 3 * N);
    m->conn = (int *)malloc(sizeof(int) * npe * T);
    memcpy(m->conn, conn, sizeof(int) * npe * T);

    /* 2D shape gradients assume counter-clockwise triangles (shape.cpp:41-42) */
    if (dim == 2) {
        for (int e = 0; e < T; ++e) {
            P3 a = node_pt(m, conn[3 * e]), b = node_pt(m, conn[3 * e + 1]), c = node_pt(m, conn[3 * e + 2]);
            if (cross(sub(b, a), sub(c, a)).z <= 0) { *status = -2; }
        }
        if (*status) { orc_mesh_free(m); return NULL; }
    }

    /* edges: first insertion fixes front/back; ids sorted by (min,max) */
    I4 *er = (I4 *)malloc(sizeof(I4) * (size_t)T * ned);
    for (int e = 0; e < T; ++e)
        for (int j = 0; j < ned; ++j) {
            int n0 = conn[e * npe + (dim == 2 ? TRI_EDGE[j][0] : TET_EDGE[j][0])];
            int n1 = conn[e * npe + (dim == 2 ? TRI_EDGE[j][1] : TET_EDGE[j][1])];
            I4 r = {n0 < n1 ? n0 : n1, n0 < n1 ? n1 : n0, e * ned + j, n0 == (n0 < n1 ? n0 : n1) ? 0 : 1};
            er[e * ned + j] = r;
        }
    qsort(er, (size_t)T * ned, sizeof(I4), cmp_i4);
    m->elem_edge = (int *)malloc(sizeof(int) * T * ned);
    int NE = 0;
    for (size_t i = 0; i < (size_t)T * ned; ++i)
        if (i == 0 || er[i].a != er[i - 1].a || er[i].b != er[i - 1].b) ++NE;
    m->NE = NE;
    m->e_front = (int *)malloc(sizeof(int) * NE);
    m->e_back = (int *)malloc(sizeof(int) * NE);
    {
        int id = -1;
        for (size_t i = 0; i < (size_t)T * ned; ++i) {
            if (i == 0 || er[i].a != er[i - 1].a || er[i].b != er[i - 1].b) {
                ++id;  /* the lowest (elem*ned+j) of a key sorts first: first insertion */
                if (er[i].d == 0) { m->e_front[id] = er[i].a; m->e_back[id] = er[i].b; }
                else { m->e_front[id] = er[i].b; m->e_back[id] = er[i].a; }
            }
            m->elem_edge[er[i].c] = id;
        }
    }
    free(er);

    /* element faces: first insertion fixes node order, edge order and tag */
    int nfn = dim == 2 ? 2 : 3;
    I4 *fr = (I4 *)malloc(sizeof(I4) * (size_t)T * nfe);
    for (int e = 0; e < T; ++e)
        for (int j = 0; j < nfe; ++j) {
            int k[3] = {0, 0, -1};
            for (int q = 0; q < nfn; ++q)
                k[q] = conn[e * npe + (dim == 2 ? TRI_FACE_NODES[j][q] : TET_FACE_NODES[j][q])];
            qsort(k, nfn, sizeof(int), cmp_int);
            I4 r = {k[0], k[1], nfn == 3 ? k[2] : -1, e * nfe + j};
            fr[e * nfe + j] = r;
        }
    qsort(fr, (size_t)T * nfe, sizeof(I4), cmp_i4);
    m->elem_face_src = (int *)malloc(sizeof(int) * T * nfe); /* (elem*nfe+j) of first insertion */
    m->face_tag_e = (int *)malloc(sizeof(int) * T * nfe);
    for (size_t i = 0; i < (size_t)T * nfe;) {
        size_t j = i;
        while (j < (size_t)T * nfe && fr[j].a == fr[i].a && fr[j].b == fr[i].b && fr[j].c == fr[i].c) ++j;
        int src = fr[i].d;
        for (size_t q = i; q < j; ++q) {
            m->elem_face_src[fr[q].d] = src;
            m->face_tag_e[fr[q].d] = eface_tag[src];
        }
        i = j;
    }
    free(fr);

    /* element centroids (element_inline.h:176-181) */
    P3 *cent = (P3 *)malloc(sizeof(P3) * T);
    for (int e = 0; e < T; ++e) {
        P3 c = p3(0, 0, 0);
        for (int q = 0; q < npe; ++q) c = add(c, node_pt(m, conn[e * npe + q]));
        cent[e] = dvd(c, (double)npe);
    }

    m->scv_vol = (double *)calloc((size_t)T * npe, sizeof(double));
    int NFI = T * ned;
    m->NFI = NFI;
    /* interior CV faces in creation order (element, local edge) */
    CvFace *cf = (CvFace *)calloc(NFI, sizeof(CvFace));

#define EDGE_MID(eid) dvd(add(node_pt(m, m->e_front[eid]), node_pt(m, m->e_back[eid])), 2.0)
    for (int e = 0; e < T; ++e) {
        const int *en = conn + e * npe;
        for (int j = 0; j < ned; ++j) {
            int eid = m->elem_edge[e * ned + j];
            int fn = m->e_front[eid], bn = m->e_back[eid];
            int lf = -1, lb = -1;
            for (int q = 0; q < npe; ++q) { if (en[q] == fn) lf = q; if (en[q] == bn) lb = q; }
            Shape s;
            P3 mid = EDGE_MID(eid);
            if (dim == 2) {
                s.np = 2; s.p[0] = mid; s.p[1] = cent[e];
                /* mesh.cpp:980-983 */
                m->scv_vol[e * npe + lf] += triangle_volume(cent[e], mid, node_pt(m, fn));
                m->scv_vol[e * npe + lb] += triangle_volume(cent[e], mid, node_pt(m, bn));
            } else {
                /* faces_from_edge: first and last element face holding the edge (mesh.cpp:1328-1363) */
                int f1 = -1, f2 = -1;
                for (int f = 0; f < 4 && f1 < 0; ++f) {
                    int src = m->elem_face_src[e * 4 + f];
                    int se = src / 4, sf = src % 4;
                    for (int q = 0; q < 3; ++q)
                        if (m->elem_edge[se * 6 + TET_FACE_EDGES[sf][q]] == eid) { f1 = f; break; }
                }
                for (int f = 3; f >= 0 && f2 < 0; --f) {
                    int src = m->elem_face_src[e * 4 + f];
                    int se = src / 4, sf = src % 4;
                    for (int q = 0; q < 3; ++q)
                        if (m->elem_edge[se * 6 + TET_FACE_EDGES[sf][q]] == eid) { f2 = f; break; }
                }
                P3 c1 = orc_face_centroid(m, m->elem_face_src[e * 4 + f1]);
                P3 c2 = orc_face_centroid(m, m->elem_face_src[e * 4 + f2]);
                s.np = 4; s.p[0] = mid; s.p[1] = c1; s.p[2] = cent[e]; s.p[3] = c2;
                /* back faces (mesh.cpp:1459-1486) and SCV pyramids (mesh.cpp:880-890) */
                for (int side = 0; side < 2; ++side) {
                    int node = side == 0 ? fn : bn;
                    int ln = side == 0 ? lf : lb;
                    /* other_edges_for_node (mesh.cpp:1413-1447) */
                    int i1 = -1, i2 = -1;
                    for (int q = 0; q < 6; ++q) {
                        int g = m->elem_edge[e * 6 + q];
                        if (g != eid && (m->e_front[g] == node || m->e_back[g] == node)) { i1 = q; break; }
                    }
                    for (int q = 5; q >= 0; --q) {
                        int g = m->elem_edge[e * 6 + q];
                        if (g != eid && (m->e_front[g] == node || m->e_back[g] == node)) { i2 = q; break; }
                    }
                    /* face_from_edges (mesh.cpp:1366-1391) */
                    int g1 = m->elem_edge[e * 6 + i1], g2 = m->elem_edge[e * 6 + i2];
                    int fi = 0, found1 = 0, found2 = 0;
                    while (fi < 4 && !(found1 && found2)) {
                        found1 = found2 = 0;
                        int src = m->elem_face_src[e * 4 + fi];
                        int se = src / 4, sf = src % 4;
                        int jj = 0;
                        while (jj < 3 && !found1) {
                            int g = m->elem_edge[se * 6 + TET_FACE_EDGES[sf][jj]];
                            found1 = (g == g1 || g == g2); ++jj;
                        }
                        while (jj < 3 && !found2) {
                            int g = m->elem_edge[se * 6 + TET_FACE_EDGES[sf][jj]];
                            found2 = (g == g1 || g == g2); ++jj;
                        }
                        ++fi;
                    }
                    int bf = fi - 1;
                    Shape bq;
                    bq.np = 4;
                    bq.p[0] = node_pt(m, node);
                    bq.p[1] = EDGE_MID(g1);
                    bq.p[2] = orc_face_centroid(m, m->elem_face_src[e * 4 + bf]);
                    bq.p[3] = EDGE_MID(g2);
                    P3 apex = dvd(add(cent[e], node_pt(m, node)), 2.0);
                    m->scv_vol[e * npe + ln] += pyramid_volume(&s, apex);
                    m->scv_vol[e * npe + ln] += pyramid_volume(&bq, apex);
                }
            }
            /* CVFace ctor: orient back -> front (cvface_inline.h:98-120) */
            CvFace *c = &cf[e * ned + j];
            c->elem = e; c->local = j; c->edge = eid; c->front = fn; c->back = bn; c->tag = 0;
            c->area = shape_area(&s);
            P3 n = shape_normal(&s);
            P3 outwards = sub(node_pt(m, fn), node_pt(m, bn));
            if (dot(n, outwards) < 0.0) { shape_reverse(&s); n = scl(n, -1.0); }
            c->nx = n.x; c->ny = n.y; c->nz = n.z;
        }
    }

    /* boundary CV faces (mesh.cpp:1046-1154), creation order kept */
    int NFB = 0;
    for (int e = 0; e < T; ++e)
        for (int j = 0; j < nfe; ++j)
            if (m->face_tag_e[e * nfe + j]) NFB += nfn;
    m->NFB = NFB;
    m->cf = (CvFace *)calloc((size_t)NFI + NFB, sizeof(CvFace));
    /* sort interior faces by (edge, element) (mesh.cpp:802; cvface_inline.h:85-95) */
    {
        I4 *key = (I4 *)malloc(sizeof(I4) * NFI);
        for (int i = 0; i < NFI; ++i) { I4 r = {cf[i].edge, cf[i].elem, i, 0}; key[i] = r; }
        qsort(key, NFI, sizeof(I4), cmp_i4);
        m->cf_of_local = (int *)malloc(sizeof(int) * NFI);
        for (int i = 0; i < NFI; ++i) {
            m->cf[i] = cf[key[i].c];
            m->cf_of_local[key[i].c] = i;
        }
        free(key);
    }
    free(cf);
    {
        int k = NFI;
        for (int e = 0; e < T; ++e)
            for (int j = 0; j < nfe; ++j) {
                int tag = m->face_tag_e[e * nfe + j];
                if (!tag) continue;
                int src = m->elem_face_src[e * nfe + j];
                int se = src / nfe, sf = src % nfe;
                P3 fc = orc_face_centroid(m, src);
                for (int q = 0; q < nfn; ++q) {
                    int node = conn[se * npe + (dim == 2 ? TRI_FACE_NODES[sf][q] : TET_FACE_NODES[sf][q])];
                    Shape s;
                    if (dim == 2) {
                        s.np = 2; s.p[0] = node_pt(m, node); s.p[1] = fc;
                    } else {
                        /* edges_from_node over the face's edges (mesh.cpp:1270-1299) */
                        int a = -1, b = -1;
                        for (int r = 0; r < 3; ++r) {
                            int g = m->elem_edge[se * 6 + TET_FACE_EDGES[sf][r]];
                            if (m->e_front[g] == node || m->e_back[g] == node) { a = g; break; }
                        }
                        for (int r = 2; r >= 0; --r) {
                            int g = m->elem_edge[se * 6 + TET_FACE_EDGES[sf][r]];
                            if (m->e_front[g] == node || m->e_back[g] == node) { b = g; break; }
                        }
                        s.np = 4; s.p[0] = node_pt(m, node); s.p[1] = EDGE_MID(a); s.p[2] = fc; s.p[3] = EDGE_MID(b);
                    }
                    P3 outwards = sub(shape_centroid(&s), cent[e]);
                    if (dot(shape_normal(&s), outwards) < 0.0) shape_reverse(&s);
                    CvFace *c = &m->cf[k];
                    c->elem = e; c->local = -1; c->edge = -1; c->front = -1; c->back = node; c->tag = tag;
                    c->area = shape_area(&s);
                    P3 n = shape_normal(&s);
                    c->nx = n.x; c->ny = n.y; c->nz = n.z;
                    ++k;
                }
            }
    }

    /* CV volumes: SCVs summed in element order (volume_inline.h:51-54) */
    m->V = (double *)calloc(N, sizeof(double));
    for (int e = 0; e < T; ++e)
        for (int q = 0; q < npe; ++q) m->V[conn[e * npe + q]] += m->scv_vol[e * npe + q];

    /* shape gradients per element (shape.cpp:22-189), one per CV face row */
    m->grad = (double *)malloc(sizeof(double) * (size_t)T * npe * 3);
    for (int e = 0; e < T; ++e) {
        const int *en = conn + e * npe;
        double *g = m->grad + (size_t)e * npe * 3;
        if (dim == 2) {
            double x0 = m->xyz[3 * en[0]], x1 = m->xyz[3 * en[1]], x2 = m->xyz[3 * en[2]];
            double y0 = m->xyz[3 * en[0] + 1], y1 = m->xyz[3 * en[1] + 1], y2 = m->xyz[3 * en[2] + 1];
            double A = 0;
            for (int q = 0; q < 3; ++q) A += m->scv_vol[e * 3 + q];
            double dx[3] = {(y1 - y2) / (2 * A), (y2 - y0) / (2 * A), (y0 - y1) / (2 * A)};
            double dy[3] = {(x2 - x1) / (2 * A), (x0 - x2) / (2 * A), (x1 - x0) / (2 * A)};
            for (int q = 0; q < 3; ++q) { g[3 * q] = dx[q]; g[3 * q + 1] = dy[q]; g[3 * q + 2] = 0; }
        } else {
            double Nxi[4] = {-1, 1, 0, 0}, Net[4] = {-1, 0, 1, 0}, Nze[4] = {-1, 0, 0, 1};
            double x_xi = 0, x_et = 0, x_ze = 0, y_xi = 0, y_et = 0, y_ze = 0, z_xi = 0, z_et = 0, z_ze = 0;
            for (int q = 0; q < 4; ++q) {
                P3 p = node_pt(m, en[q]);
                x_xi += p.x * Nxi[q]; x_et += p.x * Net[q]; x_ze += p.x * Nze[q];
                y_xi += p.y * Nxi[q]; y_et += p.y * Net[q]; y_ze += p.y * Nze[q];
                z_xi += p.z * Nxi[q]; z_et += p.z * Net[q]; z_ze += p.z * Nze[q];
            }
            double del = orc_det3(x_xi, y_xi, z_xi, x_et, y_et, z_et, x_ze, y_ze, z_ze);
            for (int q = 0; q < 4; ++q) {
                double dx = orc_det3(Nxi[q], y_xi, z_xi, Net[q], y_et, z_et, Nze[q], y_ze, z_ze);
                double dy = orc_det3(x_xi, Nxi[q], z_xi, x_et, Net[q], z_et, x_ze, Nze[q], z_ze);
                double dz = orc_det3(x_xi, y_xi, Nxi[q], x_et, y_et, Net[q], x_ze, y_ze, Nze[q]);
                g[3 * q] = dx / del; g[3 * q + 1] = dy / del; g[3 * q + 2] = dz / del;
            }
        }
    }

    /* shape-gradient CSR over interior faces, columns sorted by node id (fvmpor.h:984-1018) */
    m->sh_ia = (int *)malloc(sizeof(int) * (NFI + 1));
    m->sh_ja = (int *)malloc(sizeof(int) * (size_t)NFI * npe);
    m->sh_dx = (double *)malloc(sizeof(double) * (size_t)NFI * npe);
    m->sh_dy = (double *)malloc(sizeof(double) * (size_t)NFI * npe);
    m->sh_dz = (double *)malloc(sizeof(double) * (size_t)NFI * npe);
    for (int i = 0; i <= NFI; ++i) m->sh_ia[i] = i * npe;
    for (int i = 0; i < NFI; ++i) {
        int e = m->cf[i].elem;
        const int *en = conn + e * npe;
        I4 idx[4];
        for (int q = 0; q < npe; ++q) { I4 r = {en[q], q, 0, 0}; idx[q] = r; }
        qsort(idx, npe, sizeof(I4), cmp_i4);
        for (int q = 0; q < npe; ++q) {
            int p = i * npe + q;
            m->sh_ja[p] = idx[q].a;
            const double *g = m->grad + (size_t)e * npe * 3 + 3 * idx[q].b;
            m->sh_dx[p] = g[0]; m->sh_dy[p] = g[1]; m->sh_dz[p] = g[2];
        }
    }

    /* flux-limiting CSR: edge -> its interior CV faces, weight 1/sum(area) (fvmpor.h:1033-1071) */
    m->fl_ia = (int *)calloc(NE + 1, sizeof(int));
    for (int i = 0; i < NFI; ++i) m->fl_ia[m->cf[i].edge + 1]++;
    for (int i = 0; i < NE; ++i) m->fl_ia[i + 1] += m->fl_ia[i];
    m->fl_ja = (int *)malloc(sizeof(int) * NFI);
    m->fl_w = (double *)malloc(sizeof(double) * NFI);
    for (int i = 0; i < NFI; ++i) m->fl_ja[i] = i; /* faces already sorted by (edge, element) */
    for (int ed = 0; ed < NE; ++ed) {
        double tot = 0.0;
        for (int p = m->fl_ia[ed]; p < m->fl_ia[ed + 1]; ++p) tot += m->cf[m->fl_ja[p]].area;
        for (int p = m->fl_ia[ed]; p < m->fl_ia[ed + 1]; ++p) m->fl_w[p] = 1.0 / tot;
    }

    /* cvflux CSR: node -> its CV faces sorted ascending, +-1/V (fvmpor.h:1079-1122) */
    m->cv_ia = (int *)calloc(N + 1, sizeof(int));
    int NF = NFI + NFB;
    for (int i = 0; i < NF; ++i) {
        if (m->cf[i].front >= 0) m->cv_ia[m->cf[i].front + 1]++;
        m->cv_ia[m->cf[i].back + 1]++;
    }
    for (int i = 0; i < N; ++i) m->cv_ia[i + 1] += m->cv_ia[i];
    int nnz = m->cv_ia[N];
    m->cv_ja = (int *)malloc(sizeof(int) * nnz);
    m->cv_w = (double *)malloc(sizeof(double) * nnz);
    {
        int *fill = (int *)malloc(sizeof(int) * N);
        memcpy(fill, m->cv_ia, sizeof(int) * N);
        for (int i = 0; i < NF; ++i) { /* ascending face id: rows come out sorted */
            if (m->cf[i].front >= 0) m->cv_ja[fill[m->cf[i].front]++] = i;
            m->cv_ja[fill[m->cf[i].back]++] = i;
        }
        free(fill);
        for (int i = 0; i < N; ++i) {
            double w = 1.0 / m->V[i];
            for (int p = m->cv_ia[i]; p < m->cv_ia[i + 1]; ++p)
                m->cv_w[p] = (i == m->cf[m->cv_ja[p]].back) ? -w : w;
        }
    }
    free(cent);
#undef EDGE_MID
    return m;
}

double orc_det3(double a11, double a12, double a13, double a21, double a22, double a23,
                double a31, double a32, double a33) {
    /* shape.cpp:10-15 */
    return a11 * a22 * a33 - a11 * a23 * a32 - a21 * a12 * a33 + a21 * a13 * a32 + a31 * a12 * a23 -
           a31 * a13 * a22;
}

/* centroid of the element face first inserted as (elem*nfe+j), in its own node order (face_inline.h:76-81) */
P3 orc_face_centroid(const OrcMesh *m, int src) {
    int nfe = nfe_of(m->dim), npe = m->npe;
    int e = src / nfe, f = src % nfe;
    P3 c = p3(0, 0, 0);
    if (m->dim == 2) {
        for (int q = 0; q < 2; ++q) c = add(c, node_pt(m, m->conn[e * npe + TRI_FACE_NODES[f][q]]));
        return dvd(c, 2.0);
    }
    for (int q = 0; q < 3; ++q) c = add(c, node_pt(m, m->conn[e * npe + TET_FACE_NODES[f][q]]));
    return dvd(c, 3.0);
}

void orc_mesh_free(OrcMesh *m) {
    if (!m) return;
    free(m->xyz); free(m->conn); free(m->elem_edge); free(m->e_front); free(m->e_back);
    free(m->elem_face_src); free(m->face_tag_e); free(m->scv_vol); free(m->cf); free(m->cf_of_local);
    free(m->V); free(m->grad); free(m->sh_ia); free(m->sh_ja); free(m->sh_dx); free(m->sh_dy);
    free(m->sh_dz); free(m->fl_ia); free(m->fl_ja); free(m->fl_w); free(m->cv_ia); free(m->cv_ja);
    free(m->cv_w);
    free(m);
}

void orc_mesh_counts(const OrcMesh *m, int *out) {
    out[0] = m->N; out[1] = m->T; out[2] = m->NE; out[3] = m->NFI; out[4] = m->NFB;
    out[5] = m->cv_ia[m->N];
}

void orc_mesh_volumes(const OrcMesh *m, double *V, double *scv) {
    memcpy(V, m->V, sizeof(double) * m->N);
    if (scv) memcpy(scv, m->scv_vol, sizeof(double) * (size_t)m->T * m->npe);
}

void orc_mesh_edges(const OrcMesh *m, int *front, int *back) {
    memcpy(front, m->e_front, sizeof(int) * m->NE);
    memcpy(back, m->e_back, sizeof(int) * m->NE);
}

/* per CV face: elem, edge, front, back, tag (5 ints) and area, nx, ny, nz (4 doubles) */
void orc_mesh_cvfaces(const OrcMesh *m, int *ints, double *dbl) {
    int NF = m->NFI + m->NFB;
    for (int i = 0; i < NF; ++i) {
        const CvFace *c = &m->cf[i];
        ints[5 * i] = c->elem; ints[5 * i + 1] = c->edge; ints[5 * i + 2] = c->front;
        ints[5 * i + 3] = c->back; ints[5 * i + 4] = c->tag;
        dbl[4 * i] = c->area; dbl[4 * i + 1] = c->nx; dbl[4 * i + 2] = c->ny; dbl[4 * i + 3] = c->nz;
    }
}

/* element shape gradients, T x npe x 3 */
void orc_mesh_gradients(const OrcMesh *m, double *g) {
    memcpy(g, m->grad, sizeof(double) * (size_t)m->T * m->npe * 3);
}

/* ---------------------------------------------------------------------------
 * Reverse Cuthill-McKee order of a graph given as CSR (ia, ja; no self loops,
 * sorted columns), as Mesh::reorder_nodes_edges / find_RCM_from_edges use it
 * (mesh.cpp:118-209, 1571-1615, Boost's cuthill_mckee_ordering, reversed).
 * Degrees count the self loop the reference adds (mesh.cpp:127-128).  seed >= 0
 * starts the first component's sweep at that node (the other components at
 * their lowest-degree node); seed < 0 starts every component at a
 * pseudo-peripheral node: from its lowest-degree node (lowest index among
 * equal degrees), repeated breadth-first level structures restarting at the
 * lowest-degree node of the last level while the eccentricity grows (George
 * and Liu).  The sweep appends each node's unvisited neighbours in
 * increasing degree, ties in adjacency order.  perm[i] = old id of new node i.
 */
static int rcm_levels(int s, const int *ia, const int *ja, const char *done, int *stamp, int mark, int *buf,
                      int *nlast, int **lastp) {
    /* buf holds the BFS in order; returns the eccentricity, the last level is buf[*lastp .. + *nlast) */
    int head = 0, tail = 0, lev_start = 0, ecc = 0;
    buf[tail++] = s;
    stamp[s] = mark;
    int lev_end = tail;
    for (;;) {
        int nstart = tail;
        for (head = lev_start; head < lev_end; ++head) {
            int v = buf[head];
            for (int k = ia[v]; k < ia[v + 1]; ++k) {
                int w = ja[k];
                if (!done[w] && stamp[w] != mark) { stamp[w] = mark; buf[tail++] = w; }
            }
        }
        if (tail == nstart) break;
        lev_start = nstart;
        lev_end = tail;
        ++ecc;
    }
    *lastp = buf + lev_start;
    *nlast = lev_end - lev_start;
    return ecc;
}

int orc_rcm(int N, const int *ia, const int *ja, int seed, int *perm) {
    int *deg = (int *)malloc(sizeof(int) * (N ? N : 1));
    char *done = (char *)calloc(N ? N : 1, 1);
    int *stamp = (int *)malloc(sizeof(int) * (N ? N : 1));
    int *order = (int *)malloc(sizeof(int) * (N ? N : 1));
    int *buf = (int *)malloc(sizeof(int) * (N ? N : 1));
    int *bydeg = (int *)malloc(sizeof(int) * (N ? N : 1));
    for (int i = 0; i < N; ++i) { deg[i] = ia[i + 1] - ia[i] + 1; stamp[i] = -1; }
    /* nodes by increasing degree, ties by index (counting sort: stable) */
    int maxd = 0;
    for (int i = 0; i < N; ++i) if (deg[i] > maxd) maxd = deg[i];
    int *cnt = (int *)calloc(maxd + 2, sizeof(int));
    for (int i = 0; i < N; ++i) cnt[deg[i] + 1]++;
    for (int d = 0; d <= maxd; ++d) cnt[d + 1] += cnt[d];
    for (int i = 0; i < N; ++i) bydeg[cnt[deg[i]]++] = i;
    free(cnt);
    int n = 0, mark = 0;
    for (int q = -1; q < N; ++q) {
        int s;
        if (q < 0) {
            if (seed < 0 || seed >= N) continue;
            s = seed;
        } else {
            s = bydeg[q];
            if (done[s]) continue;
            if (seed < 0) {
                int nlast, *last;
                int ecc = rcm_levels(s, ia, ja, done, stamp, mark++, buf, &nlast, &last);
                for (;;) {
                    int best = last[0];
                    for (int k = 1; k < nlast; ++k) {
                        int v = last[k];
                        if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
                    }
                    int e2 = rcm_levels(best, ia, ja, done, stamp, mark++, buf, &nlast, &last);
                    if (e2 <= ecc) break;
                    s = best;
                    ecc = e2;
                }
            }
        }
        /* Cuthill-McKee sweep: neighbours of each node in increasing degree (insertion sort, stable) */
        int head = n;
        order[n++] = s;
        done[s] = 1;
        while (head < n) {
            int v = order[head++];
            int n0 = n;
            for (int k = ia[v]; k < ia[v + 1]; ++k)
                if (!done[ja[k]]) { done[ja[k]] = 1; order[n++] = ja[k]; }
            for (int a = n0 + 1; a < n; ++a) {
                int x = order[a], b = a;
                while (b > n0 && deg[order[b - 1]] > deg[x]) { order[b] = order[b - 1]; --b; }
                order[b] = x;
            }
        }
    }
    for (int i = 0; i < N; ++i) perm[i] = order[N - 1 - i];
    free(deg); free(done); free(stamp); free(order); free(buf); free(bydeg);
    return 0;
}
