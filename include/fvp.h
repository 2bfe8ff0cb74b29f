/*
 * libfvp: MI355X (gfx950) CVFE Richards residual and Newton-Krylov inner loop.
 *
 * C ABI.  All array arguments are plain pointers with sizes; "device" marks
 * pointers into HIP device memory of the context's device, everything else is
 * host memory.  Every entry point returns an int status: 0 = success,
 * > 0 = recoverable failure (the IDA convention of ida_integrator.h:18-36),
 * < 0 = fatal / invalid argument.  fvp_last_error() gives the message.
 * A context is bound to one device and one HIP stream and is not thread safe.
 *
 * Node numbering: the state vectors at this boundary use the node order the
 * caller passed to fvp_create (the reference renumbers at mesh load with RCM,
 * mesh.cpp:118-209; a caller that wants that ordering passes renumbered
 * arrays).  Edges are numbered as the reference numbers them: sorted by
 * (min, max) node id, front/back fixed by the first element that lists the
 * edge (mesh.cpp:385-395, 157-166).
 */
#ifndef FVP_H
#define FVP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvp_ctx fvp_ctx;

/* Zone parameter row (per physical zone): K_xx, K_yy, K_zz, phi, alpha_vG, n_vG,
 * S_r, S_s (PhysicalZone, definitions.h:203-230; S_s is the specific-storage
 * extension, 0 reproduces the reference).  Boundary row (per boundary tag):
 * value, b, eta, qx, qy, qz (BoundaryCondition, definitions.h:105-200). */
#define FVP_ZONE_FIELDS 8
#define FVP_BC_FIELDS 6

typedef struct {
    int dim;                  /* 2 (triangles) or 3 (tetrahedra) */
    int n_nodes;
    int n_elements;
    const double *xyz;        /* n_nodes x 3 */
    const int *conn;          /* n_elements x (dim+1) */
    const int *face_tags;     /* n_elements x (dim+1), 0 = interior face */
    const int *node_tag_ptr;  /* n_nodes + 1 (CSR of boundary tags per node) */
    const int *node_tags;
    const int *elem_zone;     /* n_elements, index into zones */
    const double *elem_kmul;  /* n_elements conductivity multiplier, NULL = 1 */
    int n_zones;
    const double *zones;      /* n_zones x FVP_ZONE_FIELDS */
    int n_tags;               /* boundary tags 0 .. n_tags-1 */
    const int *bc_type;       /* n_tags: 0 none, 1, 3, 4, 6, 7, 8 */
    const double *bc;         /* n_tags x FVP_BC_FIELDS */
    double rho0;              /* reference density (Constants, definitions.h:296-326) */
} fvp_problem;

const char *fvp_last_error(void);

/* Reverse Cuthill-McKee node order of a simplex mesh (conn: n_elements x
 * nodes_per_element node ids): perm[i] = the old id of new node i.  Replaces
 * Mesh::reorder_nodes_edges / find_RCM_from_edges (mesh.cpp:118-209,
 * 1571-1615; Boost's cuthill_mckee_ordering with a pseudo-peripheral start
 * per component, reversed).  The caller renumbers its arrays with it before
 * fvp_create (edges are then numbered by sorted node ids, as mesh.cpp:157-166
 * does).  Host only. */
int fvp_rcm(int n_nodes, int n_elements, int nodes_per_element, const int *conn, int *perm);

/* Build the device-side problem: replaces mesh::Mesh construction
 * (mesh.cpp:70-80) plus VarSatPhysicsImpl::initialise_vectors and
 * initialise_shape_functions (fvmpor.h:346-644, 947-1124). */
int fvp_create(const fvp_problem *problem, int device, fvp_ctx **out);
int fvp_destroy(fvp_ctx *ctx);
/* sizes[0..5] = n_nodes, n_elements, n_edges, n_dirichlet, dim, n_zones */
int fvp_sizes(const fvp_ctx *ctx, int *sizes);
/* Run all work of this context on a caller HIP stream (hipStream_t); NULL = the context's own. */
int fvp_set_stream(fvp_ctx *ctx, void *hip_stream);
int fvp_synchronize(fvp_ctx *ctx);
/* Update boundary values per tag (n_tags x FVP_BC_FIELDS), e.g. a time series
 * evaluated by the caller (BoundaryCondition::value(t), definitions.h:148-155). */
int fvp_set_bc(fvp_ctx *ctx, const double *bc);

/* Residual F(t, psi, psi') (device vectors, n_nodes each).
 * Replaces Callback::operator() -> SolverBase::compute_residual ->
 * FVMAssembler::compute_residual -> preprocess_evaluation + residual_evaluation
 * (fvm.h:348-359, solver.h:123-133, fvm_assembler.h:62-83,
 * fvmpor_ODE_impl.cpp:62-115). */
int fvp_residual(fvp_ctx *ctx, double t, const double *psi, const double *psip, double *F);

/* Freeze upwind weights for a time step: one averaged-weight flux evaluation
 * and the edge-flux sign test (Physics::preprocess_timestep,
 * fvmpor_ODE_impl.cpp:118-142; process_spatial_weights, fvmpor.h:761-797). */
int fvp_begin_step(fvp_ctx *ctx, double t, const double *psi, const double *psip);
/* Host copies of the per-edge front weight (1 = front upwind, 0 = back,
 * 0.5 = averaging) and of the averaged edge flux of the last begin_step.
 * Either pointer may be NULL.  n_edges entries each. */
int fvp_get_upwind(fvp_ctx *ctx, double *w_front, double *edge_flux);
/* Set per-edge front weights from the host (values 0, 0.5 or 1). */
int fvp_set_upwind(fvp_ctx *ctx, const double *w_front);

/* Per-node capacity coefficient ahh (device, n_nodes) of the last residual
 * evaluation (process_derivative_coefficients, fvmpor.h:930-945). */
int fvp_capacity(fvp_ctx *ctx, double *ahh);

/* Finite-difference Jacobian-vector product of the DAE iteration matrix
 *   Jv = (F(t, psi + e v, psi' + c_j e v) - F0) / e,  e = sqrt(eps_mach) ||psi||_2 / ||v||_2
 * (IDA's DQ Jv, ida_integrator.h:233-265; increment of eem_integrator.h:323-337).
 * F0 = F(t, psi, psi').  All vectors device, n_nodes. */
int fvp_jv(fvp_ctx *ctx, double t, const double *psi, const double *psip, const double *F0,
           double c_j, const double *v, double *Jv);

/* ---------------------------------------------------------------------------
 * Sparse linear algebra (the decomp C API on the device).
 * A CSR matrix is 0-based with sorted column indices per row (Tmtx_CRS with
 * block = 0, linalg_sparse.h:21-50).  Preconditioner type tags follow
 * precon.h:8-13.  decomp's build calls return 1 on success (precon.c:151-195);
 * here 0 = success and a missing / zero pivot returns FVP_ERR_PIVOT (> 0). */
#define FVP_PRECON_NONE 0
/* One-level Schur complement (PRECON_SCHUR, precon.h:9; precon_schur,
 * schur.c:1247-1323, schur_form_local :887-981, precon_schur_apply
 * :1583-1626).  Subdomains are the blocks of FVP_BLOCK_ROWS consecutive rows
 * (fvp_precon_block_rows); an interior row couples only inside its subdomain,
 * the others are boundary rows.  B (interior) gets block ILU(0), one block per
 * subdomain; each local S_i = C_ii - F_i M_B^-1 E_i is formed dense and
 * inverted (Gauss-Jordan, partial pivoting); S's couplings between subdomains
 * are dropped (block-Jacobi on S at level 0).  At most 128 boundary rows per
 * subdomain: the build fails (< 0) beyond that. */
#define FVP_PRECON_SCHUR 1
#define FVP_PRECON_JACOBI 2
#define FVP_PRECON_ILU0 5
/* Block-Jacobi ILU(0): ILU(0) of the diagonal blocks of consecutive
 * FVP_BLOCK_ROWS rows (entries coupling blocks are dropped), factored and
 * applied one workgroup per block.  The 1-GPU form of the block
 * preconditioners decomp builds per subdomain (precon_schur, schur.c:1247-1323
 * with ILU(0) on the B_i blocks).  No counterpart tag in precon.h. */
#define FVP_PRECON_BJILU0 7
/* The same block-Jacobi ILU(0) with the factorisation and both triangular
 * solves in FP32 (the input matrix and the vectors at the boundary stay FP64).
 * No counterpart in the reference. */
#define FVP_PRECON_BJILU0_F32 8
#define FVP_BLOCK_ROWS 512
#define FVP_ERR_PIVOT 1
#define FVP_ERR_NOCONV 2

typedef struct fvp_csr fvp_csr;

/* Pattern from host arrays (ia: n+1, ja: ia[n]); values are set separately. */
int fvp_csr_create(int device, int n, const int *ia, const int *ja, fvp_csr **out);
int fvp_csr_destroy(fvp_csr *A);
int fvp_csr_set_stream(fvp_csr *A, void *hip_stream);
/* values: nnz doubles, device pointer (copied) */
int fvp_csr_set_values(fvp_csr *A, const double *values);
/* host copy of the matrix values (nnz doubles) */
int fvp_csr_get_values(fvp_csr *A, double *values_host);
/* y = alpha A x + beta y (mtx_CRS_gemv, linalg_sparse.c:1321-1348), device vectors */
int fvp_csr_gemv(fvp_csr *A, double alpha, const double *x, double beta, double *y);
/* build the preconditioner of A (precon(), precon.c:151-195; ILU0_factor,
 * ILU.c:12-115 with the diagonal of row i inverted; precon_jacobi, MR.c:143-200) */
int fvp_precon_build(fvp_csr *A, int type);
/* Rows per block of FVP_PRECON_BJILU0(_F32): a power of two in [64, 4096],
 * default FVP_BLOCK_ROWS; takes effect at the next build.  The build fails
 * (< 0) when one block's factor does not fit the workgroup's LDS. */
int fvp_precon_block_rows(fvp_csr *A, int rows);
/* Overlap of the block-Jacobi blocks: each block adds `layers` (0..2) layers of
 * graph-neighbour rows to its owned rows, factors ILU(0) on the extended block
 * (rows in global order) and keeps only its owned rows of the solve
 * (restricted additive Schwarz); 0 (default) is plain block-Jacobi.  The
 * 1-GPU form of decomp's overlapping subdomain preconditioners; takes effect
 * at the next build. */
int fvp_precon_overlap(fvp_csr *A, int layers);
/* y = M^{-1} x (precon_apply, precon.c:112-149) */
int fvp_precon_apply(fvp_csr *A, const double *x, double *y);
/* host copy of the factored values (ILU(0): L unit-lower, U with inverted
 * diagonal; not available for overlapping block-Jacobi blocks) */
int fvp_precon_values(fvp_csr *A, double *values_host);

/* GMRES parameters and statistics (Tgmres, gmres.h:15-32) */
typedef struct {
    int dim_krylov;       /* restart length m */
    int max_restarts;
    int reorthogonalise;  /* classical Gram-Schmidt: 0 one pass, 1 a second pass when the first
                             cancels more than half of ||w||^2, 2 always two passes (CGS2) */
    double tol;           /* ||b - A x|| / ||b|| */
} fvp_gmres_params;
typedef struct {
    int restarts;
    int iterations;       /* Arnoldi steps over all restarts */
    int converged;
    double residual;      /* ||b - A x|| */
    double error;         /* residual / ||b|| */
} fvp_gmres_stats;

/* Right-preconditioned restarted GMRES on A x = b with A's built
 * preconditioner (gmres(), gmres.c:26-201).  x holds the initial guess. */
int fvp_gmres_csr(fvp_csr *A, const double *b, double *x, const fvp_gmres_params *p, fvp_gmres_stats *st);
/* Restarted GMRES with deflated restarts (gmresFH, gmres.c:212-405; deflate,
 * gmres.c:527-849): before each restart the k harmonic Ritz vectors of
 * smallest magnitude (a complex pair contributes its real and imaginary
 * parts) are kept at the front of the Krylov basis, with the matching block of
 * the Hessenberg matrix.  Each cycle runs the Arnoldi process to dim_krylov
 * columns and convergence is checked on the true residual after each cycle,
 * as there.  0 <= 2k < dim_krylov; k = 0 is plain restarted GMRES. */
int fvp_gmres_csr_deflated(fvp_csr *A, const double *b, double *x, const fvp_gmres_params *p, int k,
                           fvp_gmres_stats *st);
/* fvp_gmres_csr without any allocation: the Krylov basis, work vectors,
 * reduction partials and Hessenberg data live in a static 32 MiB device array
 * (one per GPU, held until the process ends) instead of library allocations,
 * and each Arnoldi step's residual estimate is read with a synchronisation
 * instead of through pinned host memory; same kernels and same results.  No
 * counterpart in the reference: for callers that cannot allocate device or
 * pinned memory (tight serving loops, harnesses that forbid allocation).  The
 * preconditioner is built beforehand (fvp_precon_build allocates).  Limit:
 * the workspace, about n (dim_krylov + 4) + 10 k doubles, must fit 4 Mi
 * doubles; larger problems return < 0 with a message.  Every call on one GPU
 * shares the array, so calls that overlap on one GPU (other threads or
 * streams) race. */
int fvp_gmres_csr_noalloc(fvp_csr *A, const double *b, double *x, const fvp_gmres_params *p,
                          fvp_gmres_stats *st);

/* ---------------------------------------------------------------------------
 * Newton-Krylov pieces on a problem context (IDA psetup / psolve and SPGMR,
 * ida_integrator.h:478-587, preconditioner_dss.cpp:284-405). */

/* Finite-difference Jacobian of the iteration matrix, stored as
 * -(dF/dpsi + c_j dF/dpsi') like the reference (preconditioner_dss.cpp:317-357),
 * then factored with the given preconditioner type.  ewt: IDA error weights
 * (device).  Returns FVP_ERR_PIVOT if the factorisation meets a zero pivot.
 * The quotients use the reference's increments; fvp_set_jacobian_mode picks
 * how they are formed (element-local by default, see below). */
int fvp_precon_setup(fvp_ctx *ctx, double t, double c_j, double h, const double *psi, const double *psip,
                     const double *F0, const double *ewt, int type);
/* z = M^{-1} r with the last setup (device vectors) */
int fvp_precon_solve(fvp_ctx *ctx, const double *r, double *z);
/* the context's Jacobian as a CSR handle (pattern = node adjacency plus diagonal) */
int fvp_jacobian(fvp_ctx *ctx, fvp_csr **A);
/* How fvp_precon_setup forms the difference quotients.
 * FVP_JACOBIAN_COLOURED: one full residual per colour of the greedy distance-2
 * colouring, as preconditioner_dss.cpp:284-372 does.
 * FVP_JACOBIAN_ELEMENT (default): every element is evaluated at the base state
 * and once per node with that node moved by its increment, in one launch, and
 * the differences are summed into the entries.  The residual is a sum of
 * element terms plus node-local terms, so these are the same quotients up to
 * the rounding of the sums.  A context with averaged k_r always uses the
 * coloured form.  No counterpart in the reference. */
#define FVP_JACOBIAN_ELEMENT 0
#define FVP_JACOBIAN_COLOURED 1
int fvp_set_jacobian_mode(fvp_ctx *ctx, int mode);
/* rows per block / overlap layers of the context's block-Jacobi preconditioner
 * (see fvp_precon_block_rows, fvp_precon_overlap) */
int fvp_set_block_rows(fvp_ctx *ctx, int rows);
int fvp_set_overlap(fvp_ctx *ctx, int layers);
/* number of colours of the distance-2 colouring (preconditioner_dss.cpp:72-94) */
int fvp_colours(fvp_ctx *ctx, int *n_colours, int *colour_host);
/* Solve J x = b, J = dF/dpsi + c_j dF/dpsi' applied matrix-free (fvp_jv) and
 * right-preconditioned by the last fvp_precon_setup. */
int fvp_newton_gmres(fvp_ctx *ctx, double t, const double *psi, const double *psip, const double *F0, double c_j,
                     const double *b, double *x, const fvp_gmres_params *p, fvp_gmres_stats *st);

/* ---------------------------------------------------------------------------
 * Time integration: variable-step, variable-order (1..max_order) BDF with
 * modified Newton and right-preconditioned, error-weight-scaled GMRES on
 * F(t, psi, psi') = 0 -- the role SUNDIALS IDAS plays for the reference
 * (IDAIntegrator, ida_integrator.h:169-340; rtol/atol and max order 3 at
 * fvmpor_ODE.cpp:79-80,176-180).  Upwind weights are frozen at the start of
 * every step (preprocess_timestep, fvmpor_ODE_impl.cpp:118-142). */
typedef struct fvp_integrator fvp_integrator;

typedef struct {
    double rtol, atol;    /* error weights ewt = 1 / (rtol |psi| + atol) */
    int max_order;        /* 1..5 */
    double h0;            /* first step; 0 = from the initial residual */
    double h_max;         /* 0 = unlimited */
    int precon_type;      /* FVP_PRECON_BJILU0(_F32) / FVP_PRECON_ILU0 / FVP_PRECON_JACOBI / FVP_PRECON_NONE */
    int dim_krylov;       /* GMRES restart length (IDA maxl) */
    int max_restarts;
    double lin_factor;    /* linear tolerance = lin_factor * newton_tol (IDA eplifac, 0.05) */
    double newton_tol;    /* IDA epsNewt, 0.33 */
    int max_newton;       /* IDA maxcor, 4 */
    int max_steps;        /* per advance call */
    int block_rows;       /* block-Jacobi block size, 0 = FVP_BLOCK_ROWS */
    int overlap;          /* block-Jacobi overlap layers (0..2); < 0 keeps the context's setting */
    int reorthogonalise;  /* GMRES second Gram-Schmidt pass: 0 never (IDA's SPGMR makes a single
                             modified Gram-Schmidt pass), 1 when the first cancelled more than half
                             of ||w||^2, 2 always */
} fvp_integrator_params;

typedef struct {
    long steps, error_fails, newton_fails;
    long residuals;       /* residual evaluations on the context since create (incl. J v, FD Jacobian) */
    long newton_iters, gmres_iters, precon_setups;
    double t, h_last;
    int order_last;
    long learned_tried;   /* steps that evaluated a learned initial iterate */
    long learned_used;    /* ... whose first Newton step estimate beat the predictor's */
} fvp_integrator_stats;

/* psi0: device, n_nodes.  Consistent psi'(t0) = F/ahh on non-Dirichlet nodes with
 * ahh != 0, 0 elsewhere (Physics::initialise, fvmpor_ODE_impl.cpp:34-60). */
int fvp_integrator_create(fvp_ctx *ctx, const fvp_integrator_params *p, double t0, const double *psi0,
                          fvp_integrator **out);
int fvp_integrator_destroy(fvp_integrator *I);
/* One accepted step, never past t_stop (IDASolve IDA_ONE_STEP with a stop time).
 * The calling thread waits for the device by polling HIP events (the norms it
 * reads are written to pinned host memory by the kernels), so it is busy for
 * the whole step; run one host thread per context that advances at once. */
int fvp_integrator_step(fvp_integrator *I, double t_stop);
/* Step until t_out (Solver::advance(t), solver.h:176-180); psi(t_out) into psi_out (device). */
int fvp_integrator_advance(fvp_integrator *I, double t_out, double *psi_out);
/* Ensemble of independent instances stored back to back (node ranges
 * [offsets[g], offsets[g+1]), offsets[0] = 0, offsets[nseg] = n_nodes): the
 * error test and the Newton test use the maximum over instances of each
 * instance's weighted RMS norm, and the linear tolerance the smallest
 * instance's share, so every instance meets its own tolerances while all of
 * them share t, h and the order.  No counterpart in the reference (one
 * problem per run); nseg <= 1 restores the single-problem norms. */
int fvp_integrator_set_segments(fvp_integrator *I, int nseg, const int *offsets);
/* current state (device pointers owned by the integrator) */
int fvp_integrator_state(fvp_integrator *I, double *t, const double **psi, const double **psip);
int fvp_integrator_get_stats(const fvp_integrator *I, fvp_integrator_stats *st);

/* ---------------------------------------------------------------------------
 * Launch shape of the residual.  Nodes are processed in tiles of tile_nodes
 * consecutive ids (a power of two in [64, 4096], default 512), one workgroup
 * per tile; xcd_map = 1 deals consecutive tiles to one XCD (shared L2 for
 * their halos).  Variants: FVP_RESIDUAL_OWNER evaluates each element once, in
 * the tile of its lowest node, with LDS sums for owned nodes and export slots
 * for the others, which a node pass gathers while it finalises (default; no
 * global atomics); FVP_RESIDUAL_WINDOW
 * stages every tile's node window in LDS and evaluates each element once per
 * tile it touches, with no global atomics and a fused finalisation (needs the
 * window to fit the LDS).  fvp_autotune times every variant / tile size / XCD
 * grouping at the state (psi, psi') and keeps the fastest; it refreshes the
 * upwind weights from that state (fvp_begin_step).  No counterpart in the
 * reference. */
#define FVP_RESIDUAL_OWNER 0
#define FVP_RESIDUAL_WINDOW 1
int fvp_set_residual_launch(fvp_ctx *ctx, int variant, int tile_nodes, int xcd_map);
int fvp_autotune(fvp_ctx *ctx, double t, const double *psi, const double *psip, int *variant, int *tile_nodes,
                 int *xcd_map, double *best_us);

/* ---------------------------------------------------------------------------
 * Live kernel timing with HIP events on the context's stream (the reference's
 * util::TimerCuda, timer_cuda.h:11-55).  While enabled every residual
 * evaluation records its element-flux kernel (events attached to the kernel's
 * dispatch) and the whole residual; kind 0 = element kernel of F, 1 = whole
 * residual (F or J v), 2 = preconditioner solve, 3 = element kernel of J v. */
int fvp_profile(fvp_ctx *ctx, int enable);
int fvp_profile_get(fvp_ctx *ctx, int kind, double *total_ms, long *count);

/* ---------------------------------------------------------------------------
 * Learned initial iterate: a message-passing network on the node-adjacency
 * graph of the mesh (CSR ia/ja without the diagonal, host arrays), 16 input
 * features per node, `layers` residual layers of width 64, one output per
 * node.  BF16 weights and activations with FP32 accumulation on the matrix
 * cores; training by the mean squared error against per-node targets and Adam
 * (beta 0.9 / 0.999, eps 1e-8) on FP32 master weights.  Parameters are one flat
 * FP32 array: We (16 x 64, row major), be (64), per layer W1 (64 x 64), W2
 * (64 x 64), b (64), then Wd (64), bd (1).  x (N x 16), y and t (N) are FP32
 * device arrays.  All work runs on the model's stream (its own, or the one
 * given to fvp_mpnn_set_stream).  No counterpart in the reference: the
 * reference starts Newton from the BDF predictor only (fvmpor_ODE.cpp:120-200). */
typedef struct fvp_mpnn fvp_mpnn;
int fvp_mpnn_create(int device, int n_nodes, const int *ia, const int *ja, int layers, unsigned long long seed,
                    fvp_mpnn **out);
int fvp_mpnn_destroy(fvp_mpnn *m);
long fvp_mpnn_num_params(const fvp_mpnn *m);
int fvp_mpnn_set_stream(fvp_mpnn *m, void *stream);
int fvp_mpnn_get_params(fvp_mpnn *m, float *host);
int fvp_mpnn_set_params(fvp_mpnn *m, const float *host);
int fvp_mpnn_get_grads(fvp_mpnn *m, float *host);
int fvp_mpnn_forward(fvp_mpnn *m, const float *x, float *y);
int fvp_mpnn_loss_grad(fvp_mpnn *m, const float *x, const float *t, double *loss);
int fvp_mpnn_adam(fvp_mpnn *m, float lr);
int fvp_mpnn_synchronize(fvp_mpnn *m);

/* Learned Newton initial iterate for the integrator.  Every step attempt
 * builds the node features x (N x 16, FP32) as
 *   x[:, 0] = psi_n, x[:, 1] = asinh(ewt * (psi_pred - psi_n)),
 *   x[:, 2] = log10(h), x[:, 3] = BDF order / 3,
 *   x[:, 4] = asinh(jac), jac = ewt * G / (c_j ahh): the Newton step of the
 *             iteration matrix's capacity diagonal (G the residual at the
 *             predictor, ahh the capacity; 0 where ahh = 0),
 *   x[:, 5:16] = the caller's static columns (static_features, device N x 16,
 *                its columns 0..4 ignored; null = zeros),
 * runs the network to y (the Newton correction in error-weight units) and
 * proposes psi_pred + c / ewt with c = y limited to |c| <= trust |jac|, of the
 * sign of jac, and 0 where psi_pred > -band (the near-saturated band, where
 * the van Genuchten-Mualem conductivity has an unbounded derivative) and on
 * Dirichlet rows.  Judge and fallback: the residual is evaluated at the
 * proposal and Newton starts from whichever of predictor and proposal has the
 * smaller first Newton step estimate, the preconditioned residual P G in the
 * error-weighted RMS norm (the weighted residual without a preconditioner);
 * the local error estimate still measures the corrector against the
 * predictor.  Back-off: after three tries in a row whose proposal did not at
 * least halve that estimate, the network sits out 4 step attempts, then 8,
 * ... up to 64 (a try that halves it resets this); learned_tried counts the
 * attempts that ran it.  m = null switches it off.  The network must be built on this
 * context's node count; it runs on the context's stream from here on.
 * Defaults of the bindings: trust 2, band 0.5 (head units). */
int fvp_integrator_set_learned(fvp_integrator *I, fvp_mpnn *m, const float *static_features, double trust,
                               double band);
/* Training-pair recording: on every step attempt the integrator writes the
 * features above (static columns from static_features, null = zeros) into
 * features (device N x 16) and, once Newton has converged, the target
 * ewt * (psi_corrected - psi_pred) into targets (device N, FP32); after
 * fvp_integrator_step returns they hold the accepted step's pair.  Null
 * features / targets switch recording off. */
int fvp_integrator_set_recording(fvp_integrator *I, const float *static_features, float *features, float *targets);

#ifdef __cplusplus
}
#endif
#endif
