// Internal declarations of libfvp (context layout, error helpers, kernel launchers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fvp.h"

namespace fvp {

void set_error(const std::string &msg);

constexpr int kMaxZones = 16;
constexpr int kMaxTags = 1 << 16;
constexpr int kTileNodes = 512;  // default nodes per tile of the tiled residual kernel

// Zone parameters passed by value to kernels: K_xx, K_yy, K_zz, phi, alpha, n, m, S_r, S_s.
struct ZoneTable {
    int nz;
    double p[kMaxZones][10];
};
enum { ZK_X = 0, ZK_Y, ZK_Z, Z_PHI, Z_ALPHA, Z_N, Z_M, Z_SR, Z_SS };

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count);
    int upload(const T *h, size_t count);
    void release();
};

}  // namespace fvp

struct fvp_jacdata;

// profiling kinds: 0 residual element kernel, 1 whole residual, 2 preconditioner
// solve, 3 J v element kernel
constexpr int kProfKinds = 4;

struct fvp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int dim = 0, N = 0, T = 0, NE = 0, npe = 0, ned = 0, nz = 0, ntags = 0;
    double rho0 = 1000.0;
    fvp::ZoneTable zt{};
    std::vector<int> bc_type;
    std::vector<double> bc;  // ntags x FVP_BC_FIELDS

    // element data (device): connectivity (int4, 2D pads the 4th slot with -1),
    // zone index, conductivity multiplier, upwind bits, front-orientation bits
    fvp::DevBuf<int4> conn;
    fvp::DevBuf<uint8_t> ezone, emask, eorient;
    fvp::DevBuf<double> kmul;
    // node data
    fvp::DevBuf<double> xyz;     // N x 3
    fvp::DevBuf<double> invV;    // 1 / CV volume
    fvp::DevBuf<uint8_t> zmask;  // bit z: node touches zone z
    fvp::DevBuf<double> zw;      // nz x N SCV-volume fraction of zone z in CV i
    fvp::DevBuf<double> kr;      // nz x N relative permeability per (zone, node)
    fvp::DevBuf<double> cap;     // N capacity coefficient ahh
    fvp::DevBuf<double> nrec;    // N x 4 node records x, y, z, head (psi + sigma v in J v) of the last residual
    fvp::DevBuf<double> acc;     // N flux accumulator
    // boundary flux: nodes with a flux condition and their rho0 * sum(q.n A)
    int nb = 0;
    fvp::DevBuf<int> bnode;
    fvp::DevBuf<double> bval;
    std::vector<int> bnode_h;
    // per boundary node, the tags of its boundary CV faces (ascending) with their
    // summed area and summed area^2 * unit normal: CSR over the node slots
    std::vector<int> bent_ptr, bent_tag;
    std::vector<double> bent_area, bent_nrm;
    // Dirichlet rows
    int jac_mode = FVP_JACOBIAN_ELEMENT;  // FD Jacobian of fvp_precon_setup
    int ndir = 0;
    fvp::DevBuf<int> dnode;
    fvp::DevBuf<double> dval;
    std::vector<int> dtag_h;       // boundary tag that sets each Dirichlet row
    std::vector<double> delev_h;   // elevation of each Dirichlet node (y in 2D, z in 3D)
    // node tiles of the residual kernel (see fvp_setup.cpp)
    int ntiles = 0, tile_nodes = fvp::kTileNodes, xcd_map = 1;  // tile size, XCD-grouped tile order
    int rvariant = FVP_RESIDUAL_OWNER;  // residual kernel variant
    long tile_elems = 0;            // element visits per residual (primary + halo)
    fvp::DevBuf<int> t_eptr, t_hptr, t_hlist, t_bptr, t_dptr;
    // element-owner variant: per element its first export (element order)
    // relative to the tile's (t_xptr); per export its slot (x_ilist), slots
    // grouped by target node (x_iptr: CSR over nodes); the slot values
    fvp::DevBuf<unsigned short> t_xoff;
    fvp::DevBuf<int> t_xptr, x_iptr, x_ilist, x_tgt;  // x_tgt: target node of each slot
    fvp::DevBuf<double> xf;
    long nexport = 0;
    fvp::DevBuf<int> t_wptr, t_whalo, t_vptr, t_tz;  // window halo nodes, visits, zones per tile
    fvp::DevBuf<unsigned short> t_lconn;             // 4 window-local node indices per visit
    int win_max = 0, ntz_max = 1;
    size_t tile_lds = 0;
    // time-step upwinding
    int averaging = 1;             // 1: all edge weights 0.5
    fvp::DevBuf<int> eedge;        // T x ned global edge of each local edge
    fvp::DevBuf<int> slot_ptr;     // NE + 1
    fvp::DevBuf<int> slot;         // T x ned (e*ned + j) sorted by (edge, element)
    fvp::DevBuf<double> qface;     // T x ned averaged q.n per (element, local edge), back->front
    fvp::DevBuf<double> ea;        // NE: summed area of the edge's CV faces (geometry, set-up)
    fvp::DevBuf<double> eflux;     // NE
    fvp::DevBuf<double> wf;        // NE front weight
    // scratch for Jv / reductions
    fvp::DevBuf<double> tmp1, tmp2, tmp3;
    fvp::DevBuf<double> red;       // reduction partials
    std::vector<double> red_h;
    fvp_jacdata *jd = nullptr;  // coloured FD Jacobian and its factorisation
    // live timing: event pairs per kind (0 element kernel, 1 residual, 2 precon solve)
    long nres = 0;  // residual evaluations issued on this context
    int prof = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev[kProfKinds];
    size_t prof_used[kProfKinds] = {};
    double prof_ms[kProfKinds] = {};
    long prof_n[kProfKinds] = {};
};

namespace fvp {
// Directional-derivative mode of the residual: the state is psi + sig v,
// psi' + c_j sig v, and the output is (F(state) - F0) / sig (sig is a device
// scalar), fused into the residual kernels.
struct JvArgs {
    const double *v = nullptr, *sig = nullptr, *F0 = nullptr;
    const double *osc = nullptr;  // optional: J v scaled row-wise by osc
    double cj = 0.;
    // optional (np > 0): the van Genuchten pass forms sigma itself from
    // ||psi||^2 (psi_n2) and np partials of ||v||^2 and writes it to sig_out
    // (== sig) for the kernels after it
    const double *psi_n2 = nullptr, *part_v = nullptr;
    double *sig_out = nullptr;
    int np = 0;
};
// J v with options for the integrator's scaled solve: psi_n2 (device scalar,
// or null) is ||psi||^2 computed once for the Newton iterate; v_n2 (or null)
// holds v_count partial sums of ||v||^2 already formed by the solve that wrote
// v (used only with psi_n2); osc (or null) scales the result row-wise.
// fvp_jv is the plain form.
int jv_impl(fvp_ctx *c, double t, const double *psi, const double *psip, const double *F0, double c_j,
            const double *v, double *Jv, const double *psi_n2, const double *osc, const double *v_n2 = nullptr,
            int v_count = 0);
// element-local FD Jacobian into J (nnz values); the arrays as in fvp_jacdata
int fd_jacobian_elem(fvp_ctx *c, double cj, const double *psi, const double *psip, const double *shift, double *ys,
                     double *kr2, double *cap2, double *contrib, const long *cptr, const int *srank, const int *ia,
                     const int *dpos, double *J, long nnz, bool fp32 = false);
// z = M^-1 (r / rdiv) with the context's preconditioner (tmp: a work vector of
// N).  With zn2 given, the block-Jacobi solves also leave partial sums of
// ||z||^2: *zn2 points at them and *zn2_count is their number (null and 0 for
// other preconditioners).
int precon_solve_div(fvp_ctx *c, const double *r, const double *rdiv, double *tmp, double *z,
                     const double **zn2 = nullptr, int *zn2_count = nullptr);
int build_tiles(fvp_ctx *c, int tile_nodes);
int edge_areas(fvp_ctx *c);
int residual_impl(fvp_ctx *c, double t, const double *psi, const double *psip, double *F, int averaging,
                  const JvArgs *jv = nullptr);
int tile_kernel_attrs(fvp_ctx *c);
int upload_bc_values(fvp_ctx *c);
int upload_dirichlet_values(fvp_ctx *c);
void jac_free(fvp_ctx *c);
// record a start / stop event of kind k on the context stream when profiling
void prof_mark(fvp_ctx *c, int k, int stop);
void prof_release(fvp_ctx *c);
// a start / stop event pair of kind k for one launch (null when not profiling)
void prof_pair(fvp_ctx *c, int k, hipEvent_t *a, hipEvent_t *b);
// node count of a network's graph
int mpnn_nodes(const fvp_mpnn *m);
}  // namespace fvp

#define FVP_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            fvp::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                \
            return -1;                                                                        \
        }                                                                                     \
    } while (0)
