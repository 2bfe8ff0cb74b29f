/*
 * TEST INFRASTRUCTURE ONLY: declarations of the CPU oracle (a plain-C
 * restatement of the reference's CVFE residual, Newton-Krylov and
 * preconditioner arithmetic).  The product library never includes this file.
 */
#ifndef FVP_ORACLE_H
#define FVP_ORACLE_H

typedef struct {
    int elem, local, edge, front, back, tag;
    double area, nx, ny, nz; /* unit normal; normal() of the reference = n * area */
} CvFace;

typedef struct {
    int dim, N, T, npe, NE, NFI, NFB;
    double *xyz;
    int *conn;
    int *elem_edge;     /* T x ned: global edge id of each local edge */
    int *e_front, *e_back;
    int *elem_face_src; /* T x nfe: (elem*nfe + j) of the first element inserting that face */
    int *face_tag_e;    /* T x nfe: boundary tag of each element face */
    double *scv_vol;    /* T x npe */
    CvFace *cf;         /* NFI interior (sorted by (edge, element)) then NFB boundary */
    int *cf_of_local;   /* (elem*ned + j) -> sorted interior face id */
    double *V;          /* N */
    double *grad;       /* T x npe x 3 */
    int *sh_ia, *sh_ja;
    double *sh_dx, *sh_dy, *sh_dz;
    int *fl_ia, *fl_ja;
    double *fl_w;
    int *cv_ia, *cv_ja;
    double *cv_w;
} OrcMesh;

OrcMesh *orc_mesh_create(int dim, int N, const double *xyz, int T, const int *conn,
                         const int *eface_tag, int *status);
void orc_mesh_free(OrcMesh *m);
double orc_det3(double, double, double, double, double, double, double, double, double);

#endif
