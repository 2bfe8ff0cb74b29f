// CVFE Richards residual on gfx950.
//
// The reference evaluates the residual as a chain of CSR SpMVs over per-CV-face
// geometry tables plus ~13 elementwise passes over 36 M face arrays
// (fvmpor_ODE_impl.cpp:62-115, fvmpor.h:646-945).  Here the CV-face geometry
// is rebuilt per element from node coordinates: on a linear simplex the
// median-dual face between nodes a and b inside element e has the area vector
//     S_ab = |e| / (d+1) * (grad N_b - grad N_a)            (oriented a -> b)
// which equals the reference's quad / segment normal times area
// (cvface_shape.h:46-78, mesh.cpp:844-1010), so the Darcy flux across it is
//     q . S_ab = |e| / (d+1) * (g_b - g_a),   g_k = grad N_k . q,
// with q = -K (grad psi + e_gravity) constant on the element (fvmpor.h:681-717).
// Kernels:
//   vg_kernel      per node: van Genuchten Sw, dSw/dpsi, k_r for every zone the
//                  CV touches, CV-averaged capacity ahh (fvmpor.h:874-945)
//   tile_lds_kernel  per node tile: gradient, 3 or 6 face fluxes and upwinded
//                  k_r of every element touching the tile, flux sums in LDS,
//                  then F = acc / V - psi' ahh, boundary fluxes, Dirichlet rows
//                  (fvmpor_ODE_impl.cpp:99-115)
//   elem_kernel    per element, atomic scatter: the averaged-weight pass of
//                  begin_step (fvmpor_ODE_impl.cpp:118-142), then finalize
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>

#include "fvp_blas.h"
#include "fvp_internal.h"

namespace fvp {

// Host twin of the reference's saturation kernel (fvmpor.h:302-343), same operation order.
// a J v row, scaled by the optional row weights of the integrator's scaled solve
__device__ __forceinline__ double jv_scaled(const JvArgs &jv, int i, double d) {
    return jv.osc ? d * jv.osc[i] : d;
}

__device__ __forceinline__ void vg_eval(double h, const double *zp, double &Sw, double &dSw, double &krw) {
    if (h >= 0.) { Sw = 1.; dSw = 0.; krw = 1.; return; }
    // The reference's operation order, with each power x^p taken as
    // exp(p ln x) on the already-rounded base (a few ulp from pow, which keeps
    // extra precision internally and costs about twice as much in FP64).
    const double nVG = zp[Z_N], mVG = zp[Z_M], S_r = zp[Z_SR];
    double d = exp(nVG * log(-zp[Z_ALPHA] * h));  // a = (alpha |h|)^n
    double s = d + 1.;                            // b
    double k = -1. / s;
    d /= s;
    s = exp(-mVG * log(s));                       // Se = b^-m
    d *= s;
    d /= h;
    d *= -(1 - S_r) * (nVG - 1);
    k += 1.;
    k = exp(mVG * log(k));                        // (1 - 1/b)^m
    k -= 1.;
    k = k * k;
    k *= sqrt(s);
    s *= (1 - S_r);
    s += S_r;
    Sw = s; dSw = d; krw = k;
}

// One node per thread; the node's zones (its mask bits) are visited in a
// per-lane loop with the zone parameters in LDS, so a wavefront pays for the
// largest zone count of its nodes rather than for every zone any of them has.
// JV with jv.np > 0: every workgroup reduces the np partials of ||v||^2 in the
// fixed order of jv_sigma_n2 (the same sigma bit for bit) and workgroup 0
// stores sigma for the finalisation, so no launch of its own is needed.
template <bool JV>
__global__ void __launch_bounds__(256) vg_kernel(int N, ZoneTable zt, double rho0, const double *__restrict__ psi,
                                                 const uint8_t *__restrict__ zmask, const double *__restrict__ zw,
                                                 double *__restrict__ kr, double *__restrict__ cap, JvArgs jv,
                                                 double *__restrict__ nrec) {
    __shared__ double zp[kMaxZones * 10];
    __shared__ double red[4];
    if (threadIdx.x < 10 * zt.nz) zp[threadIdx.x] = zt.p[threadIdx.x / 10][threadIdx.x % 10];
    // the node's own inputs are loaded first: their latency overlaps sigma's reduction
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < N;
    double h = live ? psi[i] : 0.;
    const double vi = JV && live ? jv.v[i] : 0.;
    unsigned m = live ? zmask[i] : 0u;
    double sg = 0.;
    if (JV && jv.np > 0) {
        // each lane's partials in order, four loads in flight at a time
        double b = 0.;
        const int bd = blockDim.x;
        for (int k0 = threadIdx.x; k0 < jv.np; k0 += 4 * bd) {
            double q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = k0 + u * bd < jv.np ? jv.part_v[k0 + u * bd] : 0.;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (k0 + u * bd < jv.np) b += q[u];
        }
        b = block_sum(b, red);  // ends with a barrier, so the zone table is in place
        if (threadIdx.x == 0) {
            red[0] = jv_sigma_of(jv.psi_n2[0], b);
            if (blockIdx.x == 0) jv.sig_out[0] = red[0];
        }
        __syncthreads();
        sg = red[0];
    } else {
        __syncthreads();
        if (JV) sg = jv.sig[0];
    }
    if (!live) return;
    if (JV) h += sg * vi;
    // the head of the node record the element kernel gathers (x, y, z are
    // written once at set-up)
    if (nrec) nrec[4 * (size_t)i + 3] = h;
    double phi = 0., Sw = 0., dSw = 0., Ss = 0.;
    while (m) {
        const int z = __ffs(m) - 1;
        m &= m - 1;
        const double *p = zp + 10 * z;
        const double w = zw[(size_t)z * N + i];
        double s, d, k;
        vg_eval(h, p, s, d, k);
        kr[(size_t)z * N + i] = k;
        phi += p[Z_PHI] * w;
        Sw += s * w;
        dSw += d * w;
        Ss += p[Z_SS] * w;
    }
    cap[i] = (phi * dSw + Ss * Sw) * rho0;
}

__device__ __forceinline__ void load_pt(const double *__restrict__ xyz, int i, double &x, double &y, double &z) {
    x = xyz[3 * i]; y = xyz[3 * i + 1]; z = xyz[3 * i + 2];
}

// Local edges in the element edge order of mesh.cpp:409-500: triangle edges
// (0,1) (1,2) (2,0); tetrahedron edges (0,1) (1,2) (2,0) (0,3) (1,3) (2,3).
// Compile-time functions, so that unrolled loops index registers directly.
__device__ constexpr int edge_a(int j) { return j % 3; }
__device__ constexpr int edge_b(int j) { return j < 3 ? (j + 1) % 3 : 3; }

// Face fluxes of one element, summed per element node (f[k] = net inflow into
// node k's control volume, rho0 k_r q.S over the element's CV faces).
// MODE 0: upwind k_r from the element's mask bits; MODE 1: averaged k_r
// (0.5 front + 0.5 back); MODE 2: averaged k_r, and instead of node sums the
// face fluxes (back -> front) are stored per (element, local edge) in qface
// for the edge-flux sign test.
// zk: per-zone K_xx, K_yy, K_zz in LDS.  JV: the state is psi + sig v.
// The per-element record: nodes, zone, conductivity multiplier, upwind and
// orientation bits (loaded ahead of use by the owner kernel's pipeline).
struct ElemRec {
    int4 c4;
    int z;
    double km;
    unsigned mask, orient;
};

template <int MODE>
__device__ __forceinline__ ElemRec load_elem(int e, const int4 *__restrict__ conn, const uint8_t *__restrict__ ezone,
                                             const double *__restrict__ kmul, const uint8_t *__restrict__ emask,
                                             const uint8_t *__restrict__ eorient) {
    ElemRec r;
    r.c4 = conn[e];
    r.z = ezone[e];
    r.km = kmul[e];
    r.mask = MODE == 0 ? emask[e] : 0u;
    r.orient = MODE != 0 ? eorient[e] : 0u;
    return r;
}

// t ? x : y on the bit patterns: keeps the choice between two register
// values a choice of values (a plain conditional between two elements of a
// register array is turned into an indexed load from the stack)
__device__ __forceinline__ double bit_select(unsigned t, double x, double y) {
    const unsigned long long m = 0ULL - (unsigned long long)(t & 1u);
    return __longlong_as_double((__double_as_longlong(x) & (long long)m) | (__double_as_longlong(y) & (long long)~m));
}

// One element's node values: coordinates, head (J v: psi + sig v) and the
// relative permeability of the element's zone at each node.
template <int DIM>
struct NodeVals {
    double px[DIM + 1], py[DIM + 1], pz[DIM + 1], h[DIM + 1], k[DIM + 1];
};

template <int DIM, bool JV>
__device__ __forceinline__ void load_nodes(const ElemRec &rec, int N, const double *__restrict__ xyz,
                                           const double *__restrict__ psi, const double *__restrict__ v, double sig,
                                           const double *__restrict__ kr, NodeVals<DIM> &nv) {
    const int n[4] = {rec.c4.x, rec.c4.y, rec.c4.z, rec.c4.w};
    const double *krz = kr + (size_t)rec.z * N;
#pragma unroll
    for (int k = 0; k < DIM + 1; ++k) {
        load_pt(xyz, n[k], nv.px[k], nv.py[k], nv.pz[k]);
        nv.h[k] = JV ? psi[n[k]] + sig * v[n[k]] : psi[n[k]];
        nv.k[k] = krz[n[k]];
    }
}

// load_nodes from the packed node records (x, y, z, head; two 16-byte loads per node)
template <int DIM>
__device__ __forceinline__ void load_nodes_rec(const ElemRec &rec, int N, const double *__restrict__ nrec,
                                               const double *__restrict__ kr, NodeVals<DIM> &nv) {
    const int n[4] = {rec.c4.x, rec.c4.y, rec.c4.z, rec.c4.w};
    const double *krz = kr + (size_t)rec.z * N;
#pragma unroll
    for (int k = 0; k < DIM + 1; ++k) {
        const double2 *r = reinterpret_cast<const double2 *>(nrec + 4 * (size_t)n[k]);
        const double2 a = r[0], b = r[1];
        nv.px[k] = a.x; nv.py[k] = a.y; nv.pz[k] = b.x; nv.h[k] = b.y;
        nv.k[k] = krz[n[k]];
    }
}

// The element's face fluxes from its record and node values; f[k] = net mass
// flow into node k (n[k] its node ids).
template <int DIM, int MODE>
__device__ __forceinline__ void element_fluxes_vals(int e, const ElemRec &rec, const NodeVals<DIM> &nv,
                                                    const double *zk, double rho0, double *__restrict__ qface,
                                                    int n[4], double f[4]) {
    constexpr int NPE = DIM + 1;
    constexpr int NED = DIM == 2 ? 3 : 6;
    n[0] = rec.c4.x; n[1] = rec.c4.y; n[2] = rec.c4.z; n[3] = rec.c4.w;
    const double *px = nv.px, *py = nv.py, *pz = nv.pz, *h = nv.h;
    const int z = rec.z;
    const double km = rec.km;
    double G[NPE][3];  // adjugate rows: grad N_k * det
    double det;
    if (DIM == 3) {
        double ux = px[1] - px[0], uy = py[1] - py[0], uz = pz[1] - pz[0];
        double vx = px[2] - px[0], vy = py[2] - py[0], vz = pz[2] - pz[0];
        double wx = px[3] - px[0], wy = py[3] - py[0], wz = pz[3] - pz[0];
        G[1][0] = vy * wz - vz * wy; G[1][1] = vz * wx - vx * wz; G[1][2] = vx * wy - vy * wx;
        G[2][0] = wy * uz - wz * uy; G[2][1] = wz * ux - wx * uz; G[2][2] = wx * uy - wy * ux;
        G[3][0] = uy * vz - uz * vy; G[3][1] = uz * vx - ux * vz; G[3][2] = ux * vy - uy * vx;
        det = ux * G[1][0] + uy * G[1][1] + uz * G[1][2];
    } else {
        double ux = px[1] - px[0], uy = py[1] - py[0];
        double vx = px[2] - px[0], vy = py[2] - py[0];
        G[1][0] = vy; G[1][1] = -vx; G[1][2] = 0.;
        G[2][0] = -uy; G[2][1] = ux; G[2][2] = 0.;
        det = ux * vy - uy * vx;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) G[0][d] = -(G[1][d] + G[2][d] + (DIM == 3 ? G[NPE - 1][d] : 0.));
    const double inv = 1.0 / det;
    // gradient of psi
    double gx = 0., gy = 0., gz = 0.;
#pragma unroll
    for (int k = 1; k < NPE; ++k) {
        double dh = h[k] - h[0];
        gx += dh * G[k][0]; gy += dh * G[k][1]; gz += dh * G[k][2];
    }
    gx *= inv; gy *= inv; gz *= inv;
    // Darcy flux without k_r: q = -K (grad psi + e_gravity), gravity on y (2D) / z (3D)
    double qx = gx * -(zk[3 * z] * km);
    double qy, qz;
    if (DIM == 2) { qy = (gy + 1.) * -(zk[3 * z + 1] * km); qz = 0.; }
    else { qy = gy * -(zk[3 * z + 1] * km); qz = (gz + 1.) * -(zk[3 * z + 2] * km); }
    double g[NPE];
#pragma unroll
    for (int k = 0; k < NPE; ++k) g[k] = G[k][0] * qx + G[k][1] * qy + G[k][2] * qz;
    // |e|/(d+1) * (grad N_b - grad N_a) . q  =  sign(det) / (d! (d+1)) * (g_b - g_a)
    const double sc = (det > 0 ? 1.0 : -1.0) / (DIM == 3 ? 24.0 : 6.0);
#pragma unroll
    for (int k = 0; k < 4; ++k) f[k] = 0.;
    const unsigned mask = rec.mask;
    const unsigned orient = rec.orient;
#pragma unroll
    for (int j = 0; j < NED; ++j) {
        const int a = edge_a(j);
        const int b = edge_b(j);
        double flux = sc * (g[b] - g[a]);  // flow a -> b
        double krf;
        if (MODE == 0) {
            krf = bit_select((mask >> j) & 1u, nv.k[a], nv.k[b]);
        } else {
            const unsigned a_front = (orient >> j) & 1u;
            double kf = bit_select(a_front, nv.k[a], nv.k[b]);
            double kb = bit_select(a_front, nv.k[b], nv.k[a]);
            krf = kf * 0.5;
            krf += kb * 0.5;
            if (MODE == 2) qface[(size_t)e * NED + j] = krf * (a_front ? -flux : flux);  // back -> front
        }
        double M = rho0 * (krf * flux);
        f[a] -= M;
        f[b] += M;
    }
}

template <int DIM, int MODE, bool JV = false>
__device__ __forceinline__ void element_fluxes_rec(int e, const ElemRec &rec, int N, const double *zk, double rho0,
                                                   const double *__restrict__ xyz, const double *__restrict__ psi,
                                                   const double *__restrict__ v, double sig,
                                                   const double *__restrict__ kr, double *__restrict__ qface,
                                                   int n[4], double f[4]) {
    NodeVals<DIM> nv;
    load_nodes<DIM, JV>(rec, N, xyz, psi, v, sig, kr, nv);
    element_fluxes_vals<DIM, MODE>(e, rec, nv, zk, rho0, qface, n, f);
}

template <int DIM, int MODE, bool JV = false>
__device__ __forceinline__ void element_fluxes(int e, int N, const double *zk, double rho0,
                                               const int4 *__restrict__ conn, const uint8_t *__restrict__ ezone,
                                               const double *__restrict__ kmul, const uint8_t *__restrict__ emask,
                                               const uint8_t *__restrict__ eorient,
                                               const double *__restrict__ xyz, const double *__restrict__ psi,
                                               const double *__restrict__ v, double sig,
                                               const double *__restrict__ kr, double *__restrict__ qface,
                                               int n[4], double f[4]) {
    const ElemRec rec = load_elem<MODE>(e, conn, ezone, kmul, emask, eorient);
    element_fluxes_rec<DIM, MODE, JV>(e, rec, N, zk, rho0, xyz, psi, v, sig, kr, qface, n, f);
}

// Atomic scatter variant: the averaged-weight residual (MODE 1, before the
// first begin_step) and begin_step's per-edge face sums (MODE 2).
template <int DIM, int MODE>
__global__ void __launch_bounds__(256) elem_kernel(int T, int N, ZoneTable zt, double rho0,
                                                   const int4 *__restrict__ conn, const uint8_t *__restrict__ ezone,
                                                   const double *__restrict__ kmul, const uint8_t *__restrict__ emask,
                                                   const uint8_t *__restrict__ eorient,
                                                   const double *__restrict__ xyz, const double *__restrict__ psi,
                                                   const double *__restrict__ kr, double *__restrict__ acc,
                                                   double *__restrict__ qface) {
    __shared__ double zk[3 * kMaxZones];
    if (threadIdx.x < 3 * zt.nz) zk[threadIdx.x] = zt.p[threadIdx.x / 3][ZK_X + threadIdx.x % 3];
    __syncthreads();
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    int n[4];
    double f[4];
    element_fluxes<DIM, MODE>(e, N, zk, rho0, conn, ezone, kmul, emask, eorient, xyz, psi, nullptr, 0., kr, qface,
                              n, f);
    if (MODE != 2) {
#pragma unroll
        for (int k = 0; k < DIM + 1; ++k) atomicAdd(&acc[n[k]], f[k]);
    }
}

// Export slots of an element evaluated in tile [lo, hi): -1 for owned nodes,
// else the next slot of the element's run starting at xd[x] (foreign nodes in
// local order).
template <int DIM>
__device__ __forceinline__ void owner_exports(const ElemRec &r, int lo, int hi, const int *__restrict__ xd, int x,
                                              int d[4]) {
    const int nn[4] = {r.c4.x, r.c4.y, r.c4.z, r.c4.w};
#pragma unroll
    for (int q = 0; q < DIM + 1; ++q) d[q] = (nn[q] >= lo && nn[q] < hi) ? -1 : xd[x++];
}

// Element-owner residual: a workgroup owns the nodes of one tile and evaluates
// the tile's primary elements (lowest node in the tile; every element has
// exactly one primary tile, so each element is evaluated once), gathering node
// data through the caches.  Contributions to owned nodes accumulate in LDS
// (ds_add_f64) and are stored, with the tile's boundary fluxes, to acc; those
// to nodes of other tiles are exports (the element's first: tile base +
// xoff[e], then one per foreign node in local order), each stored to its slot
// xdst[export]; slots are grouped by target node, so the finalisation reads
// each node's imports contiguously.  No global atomics.
template <int DIM, bool JV>
__global__ void __launch_bounds__(512) owner_kernel(
    int N, int ntiles, int tile_nodes, int per_xcd, ZoneTable zt, double rho0, const int4 *__restrict__ conn,
    const uint8_t *__restrict__ ezone, const double *__restrict__ kmul, const uint8_t *__restrict__ emask,
    const double *__restrict__ xyz, const double *__restrict__ nrec, const double *__restrict__ kr,
    const int *__restrict__ eptr, const int *__restrict__ bptr, const int *__restrict__ bnode,
    const double *__restrict__ bval, const unsigned short *__restrict__ xoff, const int *__restrict__ xptr,
    const int *__restrict__ xdst, double *__restrict__ xf, double *__restrict__ acc_g, JvArgs jv) {
    extern __shared__ double sm[];
    const int tile = per_xcd ? (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3) : blockIdx.x;
    if (tile >= ntiles) return;  // uniform for the whole workgroup
    // nrec: the packed node records (x, y, z, head) written by the van Genuchten
    // pass, whose head is psi + sigma v for J v; the difference quotient is
    // formed by the finalisation (xyz is read through the records)
    (void)jv;
    (void)xyz;
    const int lo = tile * tile_nodes, hi = min(N, lo + tile_nodes), nown = hi - lo;
    for (int i = threadIdx.x; i < nown; i += blockDim.x) sm[i] = 0.;
    __shared__ double zk[3 * kMaxZones];
    if (threadIdx.x < 3 * zt.nz) zk[threadIdx.x] = zt.p[threadIdx.x / 3][ZK_X + threadIdx.x % 3];
    __syncthreads();
    const int *xd = xdst + xptr[tile];
    const int e1 = eptr[tile + 1];
    const int bs = blockDim.x;
    // Two-stage software pipeline over the thread's elements: while element e
    // is evaluated, the node values and export slots of element e + bs and the
    // record of element e + 2 bs are in flight, so the gathers' latency hides
    // behind the flux arithmetic.  Look-ahead indices past the tile's range are
    // clamped to its last element (loaded, never used).
    int e = eptr[tile] + threadIdx.x;
    if (e < e1) {
        const int last = e1 - 1;
        ElemRec cur = load_elem<0>(e, conn, ezone, kmul, emask, nullptr);
        NodeVals<DIM> vc;
        load_nodes_rec<DIM>(cur, N, nrec, kr, vc);
        int dc[4];
        owner_exports<DIM>(cur, lo, hi, xd, xoff[e], dc);
        int en = min(e + bs, last);
        ElemRec nxt = load_elem<0>(en, conn, ezone, kmul, emask, nullptr);
        int xn = xoff[en];
        for (; e < e1; e += bs) {
            NodeVals<DIM> vn;
            int dn[4];
            load_nodes_rec<DIM>(nxt, N, nrec, kr, vn);
            owner_exports<DIM>(nxt, lo, hi, xd, xn, dn);
            const int e2 = min(e + 2 * bs, last);
            const ElemRec nx2 = load_elem<0>(e2, conn, ezone, kmul, emask, nullptr);
            const int x2 = xoff[e2];
            int n[4];
            double f[4];
            element_fluxes_vals<DIM, 0>(e, cur, vc, zk, rho0, nullptr, n, f);
#pragma unroll
            for (int q = 0; q < DIM + 1; ++q) {
                if (dc[q] < 0) atomicAdd(&sm[n[q] - lo], f[q]);
                else xf[dc[q]] = f[q];
            }
            cur = nxt;
            vc = vn;
#pragma unroll
            for (int q = 0; q < 4; ++q) dc[q] = dn[q];
            nxt = nx2;
            xn = x2;
        }
    }
    __syncthreads();
    for (int k = bptr[tile] + threadIdx.x; k < bptr[tile + 1]; k += blockDim.x) sm[bnode[k] - lo] -= bval[k];
    __syncthreads();
    for (int i = threadIdx.x; i < nown; i += blockDim.x) acc_g[lo + i] = sm[i];
}

// Finalisation of the element-owner residual, one workgroup per tile: the
// tile's imports are a contiguous slot range (slots are grouped by target
// node), read coalesced and added to the nodes' own sums in LDS, then
// F = acc / V - psi' ahh (JV: the difference quotient) and the Dirichlet rows.
template <bool JV>
__global__ void __launch_bounds__(1024) owner_finalize(
    int N, int ntiles, int tile_nodes, const double *__restrict__ acc, const int *__restrict__ iptr,
    const int *__restrict__ xtgt, const double *__restrict__ xf, const double *__restrict__ invV,
    const double *__restrict__ psi, const double *__restrict__ psip, const double *__restrict__ cap,
    const int *__restrict__ dptr, const int *__restrict__ dnode, const double *__restrict__ dval,
    double *__restrict__ F, JvArgs jv) {
    extern __shared__ double sm[];
    const int tile = blockIdx.x;
    const int lo = tile * tile_nodes, hi = min(N, lo + tile_nodes), nown = hi - lo;
    for (int i = threadIdx.x; i < nown; i += blockDim.x) sm[i] = acc[lo + i];
    __syncthreads();
    for (int k = iptr[lo] + threadIdx.x; k < iptr[hi]; k += blockDim.x) atomicAdd(&sm[xtgt[k] - lo], xf[k]);
    __syncthreads();
    const double sig = JV ? jv.sig[0] : 0.;
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const double a = sm[i - lo];
        if (JV) {
            const double yp = psip[i] + jv.cj * (sig * jv.v[i]);
            const double F1 = a * invV[i] - yp * cap[i];
            F[i] = jv_scaled(jv, i, (F1 - jv.F0[i]) / sig);
        } else
            F[i] = a * invV[i] - psip[i] * cap[i];
    }
    __syncthreads();
    for (int k = dptr[tile] + threadIdx.x; k < dptr[tile + 1]; k += blockDim.x) {
        const int i = dnode[k];
        if (JV) {
            const double y = psi[i] + sig * jv.v[i];
            F[i] = jv_scaled(jv, i, ((dval[k] - y) - jv.F0[i]) / sig);
        } else
            F[i] = dval[k] - psi[i];
    }
}

// LDS-staged owner-computes residual.  A workgroup owns tile_nodes nodes; it
// stages the coordinates, heads and relative permeabilities of its window (owned
// nodes plus the halo nodes its elements reach) in LDS with coalesced loads for
// the owned range and one gather per halo node, then evaluates every element
// touching the tile from LDS using the window-local connectivity (8 B per
// element visit), accumulates the flux sums of owned nodes in LDS (ds_add_f64),
// and finishes F = acc / V - psi' ahh, boundary fluxes and Dirichlet rows.
// Elements shared by several tiles are evaluated once per tile and only the
// contributions to owned nodes are kept, so no global atomics are issued; the
// tiles of one XCD are contiguous so their halo nodes share that XCD's L2.
// JV: directional-derivative mode (state psi + sig v, output (F - F0) / sig).
template <int DIM, bool JV>
__global__ void __launch_bounds__(256) tile_lds_kernel(
    int N, int ntiles, int tile_nodes, int per_xcd, int wmax, int ntzmax, ZoneTable zt, double rho0, const uint8_t *__restrict__ ezone,
    const double *__restrict__ kmul, const uint8_t *__restrict__ emask, const double *__restrict__ xyz,
    const double *__restrict__ psi, const double *__restrict__ kr, const int *__restrict__ eptr,
    const int *__restrict__ hptr, const int *__restrict__ hlist, const int *__restrict__ wptr,
    const int *__restrict__ whalo, const int *__restrict__ vptr, const ushort4 *__restrict__ lconn,
    const int *__restrict__ tzone, const double *__restrict__ invV, const double *__restrict__ psip,
    const double *__restrict__ cap, const int *__restrict__ bptr, const int *__restrict__ bnode,
    const double *__restrict__ bval, const int *__restrict__ dptr, const int *__restrict__ dnode,
    const double *__restrict__ dval, double *__restrict__ F, JvArgs jv) {
    extern __shared__ double sm[];
    // workgroups are dealt to the 8 XCDs round-robin: with per_xcd > 0 the
    // tiles of one XCD are contiguous (their halos share its L2)
    const int tile = per_xcd ? (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3) : blockIdx.x;
    if (tile >= ntiles) return;  // uniform for the whole workgroup
    const double sig = JV ? jv.sig[0] : 0.;
    const int lo = tile * tile_nodes, hi = min(N, lo + tile_nodes), nown = hi - lo;
    const int w0 = wptr[tile], win = nown + wptr[tile + 1] - w0;
    int tz[4], ntz = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        tz[q] = tzone[tile * 4 + q];
        ntz += tz[q] >= 0;
    }
    double *sx = sm, *sy = sx + wmax, *sz = sy + wmax, *sp = sz + wmax, *sk = sp + wmax;
    double *acc = sk + ntzmax * wmax;
    for (int i = threadIdx.x; i < win; i += blockDim.x) {
        const int g = i < nown ? lo + i : whalo[w0 + i - nown];
        sx[i] = xyz[3 * g];
        sy[i] = xyz[3 * g + 1];
        sz[i] = xyz[3 * g + 2];
        sp[i] = JV ? psi[g] + sig * jv.v[g] : psi[g];
        for (int q = 0; q < ntz; ++q) sk[q * wmax + i] = kr[(size_t)tz[q] * N + g];
    }
    for (int i = threadIdx.x; i < nown; i += blockDim.x) acc[i] = 0.;
    // per-zone conductivity tensor diagonal in LDS: elements of one wavefront
    // index it with their own zone, which a kernel-argument table would turn
    // into select chains over every zone
    __shared__ double zk[3 * kMaxZones];
    if (threadIdx.x < 3 * zt.nz) zk[threadIdx.x] = zt.p[threadIdx.x / 3][ZK_X + threadIdx.x % 3];
    __syncthreads();
    constexpr int NPE = DIM + 1;
    constexpr int NED = DIM == 2 ? 3 : 6;
    const int e0 = eptr[tile], np = eptr[tile + 1] - e0, h0 = hptr[tile];
    const int v0 = vptr[tile], nv = vptr[tile + 1] - v0;
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        const int e = k < np ? e0 + k : hlist[h0 + k - np];
        const ushort4 lc = lconn[v0 + k];
        const int l[4] = {lc.x, lc.y, lc.z, lc.w};
        const int z = ezone[e];
        const double km = kmul[e];
        const unsigned mask = emask[e];
        const int slot = z == tz[0] ? 0 : z == tz[1] ? 1 : z == tz[2] ? 2 : 3;
        double px[NPE], py[NPE], pz[NPE], h[NPE];
#pragma unroll
        for (int q = 0; q < NPE; ++q) {
            px[q] = sx[l[q]]; py[q] = sy[l[q]]; pz[q] = sz[l[q]]; h[q] = sp[l[q]];
        }
        double G[NPE][3];
        double det;
        if (DIM == 3) {
            double ux = px[1] - px[0], uy = py[1] - py[0], uz = pz[1] - pz[0];
            double vx = px[2] - px[0], vy = py[2] - py[0], vz = pz[2] - pz[0];
            double wx = px[3] - px[0], wy = py[3] - py[0], wz = pz[3] - pz[0];
            G[1][0] = vy * wz - vz * wy; G[1][1] = vz * wx - vx * wz; G[1][2] = vx * wy - vy * wx;
            G[2][0] = wy * uz - wz * uy; G[2][1] = wz * ux - wx * uz; G[2][2] = wx * uy - wy * ux;
            G[3][0] = uy * vz - uz * vy; G[3][1] = uz * vx - ux * vz; G[3][2] = ux * vy - uy * vx;
            det = ux * G[1][0] + uy * G[1][1] + uz * G[1][2];
        } else {
            double ux = px[1] - px[0], uy = py[1] - py[0];
            double vx = px[2] - px[0], vy = py[2] - py[0];
            G[1][0] = vy; G[1][1] = -vx; G[1][2] = 0.;
            G[2][0] = -uy; G[2][1] = ux; G[2][2] = 0.;
            det = ux * vy - uy * vx;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) G[0][d] = -(G[1][d] + G[2][d] + (DIM == 3 ? G[NPE - 1][d] : 0.));
        const double inv = 1.0 / det;
        double gx = 0., gy = 0., gz = 0.;
#pragma unroll
        for (int q = 1; q < NPE; ++q) {
            double dh = h[q] - h[0];
            gx += dh * G[q][0]; gy += dh * G[q][1]; gz += dh * G[q][2];
        }
        gx *= inv; gy *= inv; gz *= inv;
        double qx = gx * -(zk[3 * z] * km);
        double qy, qzz;
        if (DIM == 2) { qy = (gy + 1.) * -(zk[3 * z + 1] * km); qzz = 0.; }
        else { qy = gy * -(zk[3 * z + 1] * km); qzz = (gz + 1.) * -(zk[3 * z + 2] * km); }
        double g[NPE];
#pragma unroll
        for (int q = 0; q < NPE; ++q) g[q] = G[q][0] * qx + G[q][1] * qy + G[q][2] * qzz;
        const double sc = (det > 0 ? 1.0 : -1.0) / (DIM == 3 ? 24.0 : 6.0);
        double f[NPE];
#pragma unroll
        for (int q = 0; q < NPE; ++q) f[q] = 0.;
        const double *krs = sk + slot * wmax;
#pragma unroll
        for (int j = 0; j < NED; ++j) {
            const int a = edge_a(j);
            const int b = edge_b(j);
            const double flux = sc * (g[b] - g[a]);
            const double krf = krs[(mask >> j) & 1u ? l[a] : l[b]];
            const double M = rho0 * (krf * flux);
            f[a] -= M;
            f[b] += M;
        }
#pragma unroll
        for (int q = 0; q < NPE; ++q)
            if (l[q] < nown) atomicAdd(&acc[l[q]], f[q]);
    }
    __syncthreads();
    for (int k = bptr[tile] + threadIdx.x; k < bptr[tile + 1]; k += blockDim.x) acc[bnode[k] - lo] -= bval[k];
    __syncthreads();
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        if (JV) {
            const double yp = psip[i] + jv.cj * (sig * jv.v[i]);
            const double F1 = acc[i - lo] * invV[i] - yp * cap[i];
            F[i] = jv_scaled(jv, i, (F1 - jv.F0[i]) / sig);
        } else
            F[i] = acc[i - lo] * invV[i] - psip[i] * cap[i];
    }
    __syncthreads();
    for (int k = dptr[tile] + threadIdx.x; k < dptr[tile + 1]; k += blockDim.x) {
        const int i = dnode[k];
        if (JV) {
            const double y = psi[i] + sig * jv.v[i];
            F[i] = jv_scaled(jv, i, ((dval[k] - y) - jv.F0[i]) / sig);
        } else
            F[i] = dval[k] - psi[i];
    }
}

__global__ void boundary_kernel(int nb, const int *__restrict__ bnode, const double *__restrict__ bval,
                                double *__restrict__ acc) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nb) acc[bnode[k]] -= bval[k];
}

template <bool JV = false>
__global__ void finalize_kernel(int N, const double *__restrict__ acc, const double *__restrict__ invV,
                                const double *__restrict__ psip, const double *__restrict__ cap,
                                double *__restrict__ F, JvArgs jv = JvArgs()) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (JV) {
        const double sig = jv.sig[0];
        const double yp = psip[i] + jv.cj * (sig * jv.v[i]);
        const double F1 = acc[i] * invV[i] - yp * cap[i];
        F[i] = jv_scaled(jv, i, (F1 - jv.F0[i]) / sig);
    } else
        F[i] = acc[i] * invV[i] - psip[i] * cap[i];
}

template <bool JV = false>
__global__ void dirichlet_kernel(int nd, const int *__restrict__ dnode, const double *__restrict__ dval,
                                 const double *__restrict__ psi, double *__restrict__ F, JvArgs jv = JvArgs()) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nd) {
        int i = dnode[k];
        if (JV) {
            const double sig = jv.sig[0];
            const double y = psi[i] + sig * jv.v[i];
            F[i] = jv_scaled(jv, i, ((dval[k] - y) - jv.F0[i]) / sig);
        } else
            F[i] = dval[k] - psi[i];
    }
}

// edge flux = sum over the edge's CV faces of q.n / sum of their areas
// (flux_lim_matrix, fvmpor.h:1033-1071; the faces in the reference's (edge,
// element) order), then the upwind test (fvmpor.h:783-794)
// The edge's slots are taken kEdgeBatch at a time: their indices, then their
// face values, are all in flight together (an edge has about five), and the
// sum still runs in slot order.
constexpr int kEdgeBatch = 8;
__global__ void edge_flux_kernel(int NE, const int *__restrict__ slot_ptr, const int *__restrict__ slot,
                                 const double *__restrict__ qface, const double *__restrict__ ea,
                                 double *__restrict__ eflux, double *__restrict__ wf) {
    int ed = blockIdx.x * blockDim.x + threadIdx.x;
    if (ed >= NE) return;
    const double w = 1. / ea[ed];
    const int s0 = slot_ptr[ed], s1 = slot_ptr[ed + 1];
    double sum = 0.;
    for (int sb = s0; sb < s1; sb += kEdgeBatch) {
        int sl[kEdgeBatch];
        double q[kEdgeBatch];
#pragma unroll
        for (int u = 0; u < kEdgeBatch; ++u) sl[u] = slot[min(sb + u, s1 - 1)];
#pragma unroll
        for (int u = 0; u < kEdgeBatch; ++u) q[u] = qface[sl[u]];
#pragma unroll
        for (int u = 0; u < kEdgeBatch; ++u)
            if (sb + u < s1) sum += w * q[u];
    }
    eflux[ed] = sum;
    wf[ed] = sum < 0. ? 1. : 0.;
}

// CV face areas |S_ab| = |e| / (d+1) |grad N_b - grad N_a| per (element, local edge)
template <int DIM>
__global__ void face_area_kernel(int T, const int4 *__restrict__ conn, const double *__restrict__ xyz,
                                 double *__restrict__ aface) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    constexpr int NPE = DIM + 1;
    constexpr int NED = DIM == 2 ? 3 : 6;
    const int4 c4 = conn[e];
    const int n[4] = {c4.x, c4.y, c4.z, c4.w};
    double px[NPE], py[NPE], pz[NPE];
#pragma unroll
    for (int k = 0; k < NPE; ++k) load_pt(xyz, n[k], px[k], py[k], pz[k]);
    double G[NPE][3];
    if (DIM == 3) {
        double ux = px[1] - px[0], uy = py[1] - py[0], uz = pz[1] - pz[0];
        double vx = px[2] - px[0], vy = py[2] - py[0], vz = pz[2] - pz[0];
        double wx = px[3] - px[0], wy = py[3] - py[0], wz = pz[3] - pz[0];
        G[1][0] = vy * wz - vz * wy; G[1][1] = vz * wx - vx * wz; G[1][2] = vx * wy - vy * wx;
        G[2][0] = wy * uz - wz * uy; G[2][1] = wz * ux - wx * uz; G[2][2] = wx * uy - wy * ux;
        G[3][0] = uy * vz - uz * vy; G[3][1] = uz * vx - ux * vz; G[3][2] = ux * vy - uy * vx;
    } else {
        double ux = px[1] - px[0], uy = py[1] - py[0];
        double vx = px[2] - px[0], vy = py[2] - py[0];
        G[1][0] = vy; G[1][1] = -vx; G[1][2] = 0.;
        G[2][0] = -uy; G[2][1] = ux; G[2][2] = 0.;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) G[0][d] = -(G[1][d] + G[2][d] + (DIM == 3 ? G[NPE - 1][d] : 0.));
    const double sc = 1.0 / (DIM == 3 ? 24.0 : 6.0);
#pragma unroll
    for (int j = 0; j < NED; ++j) {
        const int a = edge_a(j), b = edge_b(j);
        const double sx = G[b][0] - G[a][0], sy = G[b][1] - G[a][1], sz = G[b][2] - G[a][2];
        aface[(size_t)e * NED + j] = sc * sqrt(sx * sx + sy * sy + sz * sz);
    }
}

__global__ void edge_area_kernel(int NE, const int *__restrict__ slot_ptr, const int *__restrict__ slot,
                                 const double *__restrict__ aface, double *__restrict__ ea) {
    int ed = blockIdx.x * blockDim.x + threadIdx.x;
    if (ed >= NE) return;
    double tot = 0.;
    for (int s = slot_ptr[ed]; s < slot_ptr[ed + 1]; ++s) tot += aface[slot[s]];
    ea[ed] = tot;
}

// per-element upwind bits: bit j set when the first local node of edge j is upwind
template <int DIM>
__global__ void mask_kernel(int T, const int *__restrict__ eedge, const uint8_t *__restrict__ eorient,
                            const double *__restrict__ wf, uint8_t *__restrict__ emask) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    constexpr int NED = DIM == 2 ? 3 : 6;
    unsigned o = eorient[e], m = 0;
#pragma unroll
    for (int j = 0; j < NED; ++j) {
        bool front_up = wf[eedge[(size_t)e * NED + j]] == 1.0;
        bool a_front = (o >> j) & 1u;
        if (front_up == a_front) m |= 1u << j;
    }
    emask[e] = (uint8_t)m;
}

static inline int blocks(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Allow the tile kernel's dynamic LDS beyond the default limit; fails cleanly if
// the device cannot provide it.
int tile_kernel_attrs(fvp_ctx *c) {
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess)
        lds_max = 64 * 1024;
    // the static zone table shares the workgroup's LDS with the dynamic window
    const size_t static_lds = sizeof(double) * 3 * kMaxZones;
    if (c->tile_lds + static_lds > (size_t)lds_max) {
        set_error("residual tile window needs " + std::to_string(c->tile_lds) + " B of LDS, device allows " +
                  std::to_string(lds_max));
        return -1;
    }
    // the attribute is per kernel, not per context: allow the device maximum so
    // that contexts with different windows can coexist
    const int dyn = lds_max - (int)static_lds;
    FVP_HIP(hipFuncSetAttribute((const void *)tile_lds_kernel<3, false>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    FVP_HIP(hipFuncSetAttribute((const void *)tile_lds_kernel<2, false>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    FVP_HIP(hipFuncSetAttribute((const void *)tile_lds_kernel<3, true>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    FVP_HIP(hipFuncSetAttribute((const void *)tile_lds_kernel<2, true>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    return 0;
}

template <int DIM, bool JV>
static void launch_tile(fvp_ctx *c, const double *psi, const double *psip, double *F, const JvArgs &jv) {
    const int per = (c->ntiles + 7) / 8;  // tiles per XCD: consecutive tiles share an XCD's L2
    hipStream_t s = c->stream;
    if (c->rvariant == FVP_RESIDUAL_OWNER) {
        const int N = c->N;
        // one workgroup per tile, wide enough that large tiles still fill the CUs; when
        // profiling, the events are part of the dispatch (the kernel's own duration)
        const int threads = c->tile_nodes >= 1024 ? 512 : 256;
        hipEvent_t ea, eb;
        prof_pair(c, JV ? 3 : 0, &ea, &eb);
        hipExtLaunchKernelGGL(owner_kernel<DIM, JV>, dim3(c->xcd_map ? per * 8 : c->ntiles), dim3(threads),
                              sizeof(double) * c->tile_nodes, s, ea, eb, 0, N, c->ntiles, c->tile_nodes,
                              c->xcd_map ? per : 0, c->zt, c->rho0, (const int4 *)c->conn.p,
                              (const uint8_t *)c->ezone.p, (const double *)c->kmul.p, (const uint8_t *)c->emask.p,
                              (const double *)c->xyz.p, (const double *)c->nrec.p,
                              (const double *)c->kr.p, (const int *)c->t_eptr.p,
                              (const int *)c->t_bptr.p, (const int *)c->bnode.p, (const double *)c->bval.p,
                              (const unsigned short *)c->t_xoff.p, (const int *)c->t_xptr.p,
                              (const int *)c->x_ilist.p, c->xf.p, c->acc.p, jv);
        owner_finalize<JV><<<c->ntiles, 1024, sizeof(double) * c->tile_nodes, s>>>(
            N, c->ntiles, c->tile_nodes, c->acc.p, c->x_iptr.p, c->x_tgt.p, c->xf.p, c->invV.p, psi, psip, c->cap.p,
            c->t_dptr.p, c->dnode.p, c->dval.p, F, jv);
        return;
    }
    hipEvent_t ea, eb;
    prof_pair(c, JV ? 3 : 0, &ea, &eb);
    if (ea) (void)hipEventRecord(ea, s);
    tile_lds_kernel<DIM, JV><<<c->xcd_map ? per * 8 : c->ntiles, 256, c->tile_lds, c->stream>>>(
        c->N, c->ntiles, c->tile_nodes, c->xcd_map ? per : 0, c->win_max, c->ntz_max, c->zt, c->rho0, c->ezone.p, c->kmul.p, c->emask.p, c->xyz.p,
        psi, c->kr.p, c->t_eptr.p, c->t_hptr.p, c->t_hlist.p, c->t_wptr.p, c->t_whalo.p, c->t_vptr.p,
        (const ushort4 *)c->t_lconn.p, c->t_tz.p, c->invV.p, psip, c->cap.p, c->t_bptr.p, c->bnode.p, c->bval.p,
        c->t_dptr.p, c->dnode.p, c->dval.p, F, jv);
    if (eb) (void)hipEventRecord(eb, s);
}

// Element-local finite-difference Jacobian, -(dF/dpsi + c_j dF/dpsi') with
// the coloured method's increments (shift) and upwind bits: every element is
// evaluated at the base state and once per node with that node's head moved
// by its shift (k_r from the second van Genuchten pass, kr2), and
// (f_base - f_moved) V^-1 / shift for (row node i, moved node k) is stored at
// contrib[srank[16 e + 4 i + k]], where the entries' contributions lie
// contiguous and in element order; fd_gather_t sums each entry's range (no
// atomics, so the Jacobian is repeatable bit for bit).  The
// capacity term and the Dirichlet rows are set by fd_node_kernel and
// fd_dirichlet.  The residual is a sum of element terms, so these are the
// coloured quotients up to the rounding of the sums.
template <int DIM>
__global__ void __launch_bounds__(256) fd_elem_kernel(int T, int N, ZoneTable zt, double rho0,
                                                      const int4 *__restrict__ conn, const uint8_t *__restrict__ ezone,
                                                      const double *__restrict__ kmul, const uint8_t *__restrict__ emask,
                                                      const double *__restrict__ nrec, const double *__restrict__ kr,
                                                      const double *__restrict__ kr2, const double *__restrict__ shift,
                                                      const double *__restrict__ invV,
                                                      const int *__restrict__ srank,
                                                      double *__restrict__ contrib) {
    __shared__ double zk[3 * kMaxZones];
    if (threadIdx.x < 3 * zt.nz) zk[threadIdx.x] = zt.p[threadIdx.x / 3][ZK_X + threadIdx.x % 3];
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    const ElemRec rec = load_elem<0>(e, conn, ezone, kmul, emask, nullptr);
    NodeVals<DIM> nv;
    load_nodes_rec<DIM>(rec, N, nrec, kr, nv);
    int n[4];
    double f0[4];
    element_fluxes_vals<DIM, 0>(e, rec, nv, zk, rho0, nullptr, n, f0);
    const double *k2 = kr2 + (size_t)rec.z * N;
    double w[DIM + 1];
#pragma unroll
    for (int i = 0; i < DIM + 1; ++i) w[i] = invV[n[i]];
#pragma unroll
    for (int k = 0; k < DIM + 1; ++k) {
        const double sk = shift[n[k]], hk = nv.h[k], kk = nv.k[k];
        nv.h[k] = hk + sk;
        nv.k[k] = k2[n[k]];
        int n1[4];
        double f1[4];
        element_fluxes_vals<DIM, 0>(e, rec, nv, zk, rho0, nullptr, n1, f1);
        nv.h[k] = hk;
        nv.k[k] = kk;
#pragma unroll
        for (int i = 0; i < DIM + 1; ++i) contrib[srank[16 * (size_t)e + 4 * i + k]] = (f0[i] - f1[i]) * w[i] / sk;
    }
}

// The same element-local quotients in FP32 arithmetic (the FP32
// preconditioner's input).  For a fixed k_r a face flux is linear in the
// heads, so with node k moved by s_k the difference quotient of the face's
// mass flow M = rho0 k_r flux is, exactly,
//   (M1 - M0) / s_k = rho0 [ (k_r1 - k_r0) / s_k * flux + k_r1 * dflux/dh_k ],
// k_r1 differing from k_r0 only when node k is the face's upwind node.  Each
// term is formed directly (dflux/dh_k from the adjugate, (k_r1 - k_r0) / s_k
// in FP64 from the two van Genuchten passes), so nothing cancels and FP32
// carries the quotient to its rounding; the geometry is built once per
// element, relative to its first node.  Same increments, same upwind bits,
// same contribution slots as fd_elem_kernel.
template <int DIM>
__global__ void __launch_bounds__(256) fd_elem_kernel32(int T, int N, ZoneTable zt, double rho0,
                                                        const int4 *__restrict__ conn,
                                                        const uint8_t *__restrict__ ezone,
                                                        const double *__restrict__ kmul,
                                                        const uint8_t *__restrict__ emask,
                                                        const double *__restrict__ nrec, const double *__restrict__ kr,
                                                        const double *__restrict__ kr2, const double *__restrict__ shift,
                                                        const double *__restrict__ invV, const int *__restrict__ srank,
                                                        float *__restrict__ contrib) {
    constexpr int NPE = DIM + 1;
    constexpr int NED = DIM == 2 ? 3 : 6;
    __shared__ float zk[3 * kMaxZones];
    if (threadIdx.x < 3 * zt.nz) zk[threadIdx.x] = (float)zt.p[threadIdx.x / 3][ZK_X + threadIdx.x % 3];
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    const ElemRec rec = load_elem<0>(e, conn, ezone, kmul, emask, nullptr);
    NodeVals<DIM> nv;
    load_nodes_rec<DIM>(rec, N, nrec, kr, nv);
    const int n[4] = {rec.c4.x, rec.c4.y, rec.c4.z, rec.c4.w};
    float G[NPE][3], det;
    if (DIM == 3) {
        const float ux = (float)(nv.px[1] - nv.px[0]), uy = (float)(nv.py[1] - nv.py[0]), uz = (float)(nv.pz[1] - nv.pz[0]);
        const float vx = (float)(nv.px[2] - nv.px[0]), vy = (float)(nv.py[2] - nv.py[0]), vz = (float)(nv.pz[2] - nv.pz[0]);
        const float wx = (float)(nv.px[3] - nv.px[0]), wy = (float)(nv.py[3] - nv.py[0]), wz = (float)(nv.pz[3] - nv.pz[0]);
        G[1][0] = vy * wz - vz * wy; G[1][1] = vz * wx - vx * wz; G[1][2] = vx * wy - vy * wx;
        G[2][0] = wy * uz - wz * uy; G[2][1] = wz * ux - wx * uz; G[2][2] = wx * uy - wy * ux;
        G[3][0] = uy * vz - uz * vy; G[3][1] = uz * vx - ux * vz; G[3][2] = ux * vy - uy * vx;
        det = ux * G[1][0] + uy * G[1][1] + uz * G[1][2];
    } else {
        const float ux = (float)(nv.px[1] - nv.px[0]), uy = (float)(nv.py[1] - nv.py[0]);
        const float vx = (float)(nv.px[2] - nv.px[0]), vy = (float)(nv.py[2] - nv.py[0]);
        G[1][0] = vy; G[1][1] = -vx; G[1][2] = 0.f;
        G[2][0] = -uy; G[2][1] = ux; G[2][2] = 0.f;
        det = ux * vy - uy * vx;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) G[0][d] = -(G[1][d] + G[2][d] + (DIM == 3 ? G[NPE - 1][d] : 0.f));
    const float inv = 1.f / det;
    const float km = (float)rec.km;
    const float Kx = -(zk[3 * rec.z] * km), Ky = -(zk[3 * rec.z + 1] * km), Kz = -(zk[3 * rec.z + 2] * km);
    // base state: head gradient, Darcy flux without k_r, face fluxes
    float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
    for (int k = 1; k < NPE; ++k) {
        const float dh = (float)(nv.h[k] - nv.h[0]);
        gx += dh * G[k][0]; gy += dh * G[k][1]; gz += dh * G[k][2];
    }
    gx *= inv; gy *= inv; gz *= inv;
    float qx = gx * Kx, qy, qz;
    if (DIM == 2) { qy = (gy + 1.f) * Ky; qz = 0.f; }
    else { qy = gy * Ky; qz = (gz + 1.f) * Kz; }
    float g[NPE];
#pragma unroll
    for (int k = 0; k < NPE; ++k) g[k] = G[k][0] * qx + G[k][1] * qy + G[k][2] * qz;
    const float sc = (det > 0.f ? 1.f : -1.f) / (DIM == 3 ? 24.f : 6.f);
    const float r0 = (float)rho0;
    const double *k2 = kr2 + (size_t)rec.z * N;
    float w[NPE], kb[NPE], dkr[NPE];
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        w[i] = (float)invV[n[i]];
        kb[i] = (float)nv.k[i];
    }
#pragma unroll
    for (int k = 0; k < NPE; ++k) {
        const double sk = shift[n[k]], k1 = k2[n[k]];
        dkr[k] = (float)((k1 - nv.k[k]) / sk);
        // d(grad h)/dh_k = G[k] / det, so dq/dh_k = K (G[k] / det) (no gravity term)
        const float dqx = G[k][0] * inv * Kx, dqy = G[k][1] * inv * Ky, dqz = G[k][2] * inv * Kz;
        float dg[NPE];
#pragma unroll
        for (int m = 0; m < NPE; ++m) dg[m] = G[m][0] * dqx + G[m][1] * dqy + G[m][2] * dqz;
        float d[4] = {0.f, 0.f, 0.f, 0.f};
        const float kmv = (float)k1;
#pragma unroll
        for (int j = 0; j < NED; ++j) {
            const int a = edge_a(j), b = edge_b(j);
            const unsigned up_a = (rec.mask >> j) & 1u;  // 1: node a is upwind
            const int u = up_a ? a : b;
            const float flux = sc * (g[b] - g[a]);
            const float dflux = sc * (dg[b] - dg[a]);
            const float kr1 = u == k ? kmv : (up_a ? kb[a] : kb[b]);
            const float dM = r0 * ((u == k ? dkr[k] : 0.f) * flux + kr1 * dflux);
            d[a] += dM;  // (f0 - f1) / s_k at a: f[a] -= M
            d[b] -= dM;
        }
#pragma unroll
        for (int i = 0; i < NPE; ++i) contrib[srank[16 * (size_t)e + 4 * i + k]] = d[i] * w[i];
    }
}

// J[p] = sum of entry p's contributions contrib[cptr[p] .. cptr[p + 1]) in
// element order, accumulated in FP64 (FP32 or FP64 contributions); the loads
// go out kEdgeBatch at a time so that they are in flight together.
template <typename C>
__global__ void fd_gather_t(long nnz, const long *__restrict__ cptr, const C *__restrict__ contrib,
                            double *__restrict__ J) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const long q0 = cptr[p], q1 = cptr[p + 1];
    double s = 0.;
    for (long qb = q0; qb < q1; qb += kEdgeBatch) {
        C c[kEdgeBatch];
#pragma unroll
        for (int u = 0; u < kEdgeBatch; ++u) c[u] = qb + u < q1 ? contrib[qb + u] : C(0);
#pragma unroll
        for (int u = 0; u < kEdgeBatch; ++u)
            if (qb + u < q1) s += (double)c[u];
    }
    J[p] = s;
}

// diagonal: the capacity term -psi' a_hh(psi) moved with psi_i and psi'_i + c_j shift_i
__global__ void fd_node_kernel(int N, double cj, const double *__restrict__ psip, const double *__restrict__ shift,
                               const double *__restrict__ cap, const double *__restrict__ cap2,
                               const int *__restrict__ dpos, double *__restrict__ J) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double sh = shift[i];
    J[dpos[i]] += ((psip[i] + cj * sh) * cap2[i] - psip[i] * cap[i]) / sh;
}

// Dirichlet rows F = psi_D - psi: zero off the diagonal, the quotient on it
__global__ void fd_dirichlet(int nd, const int *__restrict__ dnode, const double *__restrict__ dval,
                             const double *__restrict__ psi, const double *__restrict__ shift,
                             const int *__restrict__ ia, const int *__restrict__ dpos, double *__restrict__ J) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nd) return;
    const int i = dnode[k];
    for (int p = ia[i]; p < ia[i + 1]; ++p) J[p] = 0.;
    J[dpos[i]] = ((dval[k] - psi[i]) - (dval[k] - (psi[i] + shift[i]))) / shift[i];
}

__global__ void fd_add(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] + b[i];
}

int fd_jacobian_elem(fvp_ctx *c, double cj, const double *psi, const double *psip, const double *shift, double *ys,
                     double *kr2, double *cap2, double *contrib, const long *cptr, const int *srank, const int *ia,
                     const int *dpos, double *J, long nnz, bool fp32) {
    hipStream_t s = c->stream;
    const int N = c->N, T = c->T;
    // base state: k_r, capacity and the node records' heads
    vg_kernel<false><<<blocks(N, 256), 256, 0, s>>>(N, c->zt, c->rho0, psi, c->zmask.p, c->zw.p, c->kr.p, c->cap.p,
                                                     JvArgs(), c->nrec.p);
    // every node moved by its shift (each node is moved once, in its own column)
    fd_add<<<blocks(N, 256), 256, 0, s>>>(N, psi, shift, ys);
    vg_kernel<false><<<blocks(N, 256), 256, 0, s>>>(N, c->zt, c->rho0, ys, c->zmask.p, c->zw.p, kr2, cap2, JvArgs(),
                                                     nullptr);
    if (fp32) {
        float *c32 = reinterpret_cast<float *>(contrib);
        if (c->dim == 3)
            fd_elem_kernel32<3><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, (const int4 *)c->conn.p,
                                                                (const uint8_t *)c->ezone.p, c->kmul.p,
                                                                (const uint8_t *)c->emask.p, c->nrec.p, c->kr.p, kr2,
                                                                shift, c->invV.p, srank, c32);
        else
            fd_elem_kernel32<2><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, (const int4 *)c->conn.p,
                                                                (const uint8_t *)c->ezone.p, c->kmul.p,
                                                                (const uint8_t *)c->emask.p, c->nrec.p, c->kr.p, kr2,
                                                                shift, c->invV.p, srank, c32);
        fd_gather_t<float><<<blocks(nnz, 256), 256, 0, s>>>(nnz, cptr, c32, J);
    } else {
        if (c->dim == 3)
            fd_elem_kernel<3><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, (const int4 *)c->conn.p,
                                                              (const uint8_t *)c->ezone.p, c->kmul.p,
                                                              (const uint8_t *)c->emask.p, c->nrec.p, c->kr.p, kr2,
                                                              shift, c->invV.p, srank, contrib);
        else
            fd_elem_kernel<2><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, (const int4 *)c->conn.p,
                                                              (const uint8_t *)c->ezone.p, c->kmul.p,
                                                              (const uint8_t *)c->emask.p, c->nrec.p, c->kr.p, kr2,
                                                              shift, c->invV.p, srank, contrib);
        fd_gather_t<double><<<blocks(nnz, 256), 256, 0, s>>>(nnz, cptr, contrib, J);
    }
    fd_node_kernel<<<blocks(N, 256), 256, 0, s>>>(N, cj, psip, shift, c->cap.p, cap2, dpos, J);
    if (c->ndir)
        fd_dirichlet<<<blocks(c->ndir, 256), 256, 0, s>>>(c->ndir, c->dnode.p, c->dval.p, psi, shift, ia, dpos, J);
    FVP_HIP(hipGetLastError());
    return 0;
}

int residual_impl(fvp_ctx *c, double t, const double *psi, const double *psip, double *F, int averaging,
                  const JvArgs *jv) {
    (void)t;
    hipStream_t s = c->stream;
    const int N = c->N, T = c->T;
    if (jv && averaging) {
        set_error("internal: fused J v needs the upwinded residual");
        return -1;
    }
    ++c->nres;
    prof_mark(c, 1, 0);
    if (jv)
        vg_kernel<true><<<blocks(N, 256), 256, 0, s>>>(N, c->zt, c->rho0, psi, c->zmask.p, c->zw.p, c->kr.p,
                                                        c->cap.p, *jv, c->nrec.p);
    else
        vg_kernel<false><<<blocks(N, 256), 256, 0, s>>>(N, c->zt, c->rho0, psi, c->zmask.p, c->zw.p, c->kr.p,
                                                         c->cap.p, JvArgs(), c->rvariant == FVP_RESIDUAL_OWNER ? c->nrec.p : nullptr);
    if (!averaging) {
        // tiled paths (element fluxes with the boundary fluxes, then the finalisation)
        const JvArgs none;
        if (c->dim == 3) {
            if (jv) launch_tile<3, true>(c, psi, psip, F, *jv);
            else launch_tile<3, false>(c, psi, psip, F, none);
        } else {
            if (jv) launch_tile<2, true>(c, psi, psip, F, *jv);
            else launch_tile<2, false>(c, psi, psip, F, none);
        }
    } else {
        FVP_HIP(hipMemsetAsync(c->acc.p, 0, sizeof(double) * N, s));
        if (c->dim == 3)
            elem_kernel<3, 1><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, c->conn.p, c->ezone.p, c->kmul.p,
                                                           c->emask.p, c->eorient.p, c->xyz.p, psi, c->kr.p,
                                                           c->acc.p, nullptr);
        else
            elem_kernel<2, 1><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, c->conn.p, c->ezone.p, c->kmul.p,
                                                           c->emask.p, c->eorient.p, c->xyz.p, psi, c->kr.p,
                                                           c->acc.p, nullptr);
        if (c->nb) boundary_kernel<<<blocks(c->nb, 256), 256, 0, s>>>(c->nb, c->bnode.p, c->bval.p, c->acc.p);
        finalize_kernel<<<blocks(N, 256), 256, 0, s>>>(N, c->acc.p, c->invV.p, psip, c->cap.p, F);
        if (c->ndir) dirichlet_kernel<<<blocks(c->ndir, 256), 256, 0, s>>>(c->ndir, c->dnode.p, c->dval.p, psi, F);
    }
    prof_mark(c, 1, 1);
    FVP_HIP(hipGetLastError());
    return 0;
}

// Set-up: summed CV face area of every edge (geometry only)
int edge_areas(fvp_ctx *c) {
    DevBuf<double> aface;
    if (aface.alloc((size_t)c->T * c->ned)) return -1;
    if (c->dim == 3) face_area_kernel<3><<<blocks(c->T, 256), 256, 0, c->stream>>>(c->T, c->conn.p, c->xyz.p, aface.p);
    else face_area_kernel<2><<<blocks(c->T, 256), 256, 0, c->stream>>>(c->T, c->conn.p, c->xyz.p, aface.p);
    edge_area_kernel<<<blocks(c->NE, 256), 256, 0, c->stream>>>(c->NE, c->slot_ptr.p, c->slot.p, aface.p, c->ea.p);
    FVP_HIP(hipGetLastError());
    FVP_HIP(hipStreamSynchronize(c->stream));
    aface.release();
    return 0;
}

// Event pairs are pooled per kind and folded into running totals every 4096 pairs.
static void prof_fold(fvp_ctx *c, int k) {
    if (!c->prof_used[k]) return;
    (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < c->prof_used[k]; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->prof_ev[k][i].first, c->prof_ev[k][i].second) == hipSuccess) {
            c->prof_ms[k] += ms;
            c->prof_n[k] += 1;
        }
    }
    c->prof_used[k] = 0;
}

void prof_mark(fvp_ctx *c, int k, int stop) {
    if (!c->prof) return;
    if (!stop) {
        if (c->prof_used[k] == 4096) prof_fold(c, k);
        if (c->prof_used[k] == c->prof_ev[k].size()) {
            hipEvent_t a, b;
            if (hipEventCreateWithFlags(&a, hipEventDisableSystemFence) != hipSuccess) return;
            if (hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess) { (void)hipEventDestroy(a); return; }
            c->prof_ev[k].push_back({a, b});
        }
        (void)hipEventRecord(c->prof_ev[k][c->prof_used[k]].first, c->stream);
    } else {
        (void)hipEventRecord(c->prof_ev[k][c->prof_used[k]].second, c->stream);
        ++c->prof_used[k];
    }
}

// a start / stop event pair of kind k for one launch (both null when not profiling)
void prof_pair(fvp_ctx *c, int k, hipEvent_t *a, hipEvent_t *b) {
    *a = *b = nullptr;
    if (!c->prof) return;
    if (c->prof_used[k] == 4096) prof_fold(c, k);
    if (c->prof_used[k] == c->prof_ev[k].size()) {
        hipEvent_t e0, e1;
        if (hipEventCreateWithFlags(&e0, hipEventDisableSystemFence) != hipSuccess) return;
        if (hipEventCreateWithFlags(&e1, hipEventDisableSystemFence) != hipSuccess) { (void)hipEventDestroy(e0); return; }
        c->prof_ev[k].push_back({e0, e1});
    }
    *a = c->prof_ev[k][c->prof_used[k]].first;
    *b = c->prof_ev[k][c->prof_used[k]].second;
    ++c->prof_used[k];
}

void prof_release(fvp_ctx *c) {
    for (int k = 0; k < kProfKinds; ++k) {
        for (auto &e : c->prof_ev[k]) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        c->prof_ev[k].clear();
        c->prof_used[k] = 0;
    }
}

}  // namespace fvp

using namespace fvp;

extern "C" int fvp_residual(fvp_ctx *c, double t, const double *psi, const double *psip, double *F) {
    if (!c || !psi || !psip || !F) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    return residual_impl(c, t, psi, psip, F, c->averaging);
}

extern "C" int fvp_begin_step(fvp_ctx *c, double t, const double *psi, const double *psip) {
    if (!c || !psi || !psip) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    (void)t; (void)psip;
    const int N = c->N, T = c->T;
    // k_r at the state, then the averaged-weight face fluxes summed per edge
    vg_kernel<false><<<blocks(N, 256), 256, 0, s>>>(N, c->zt, c->rho0, psi, c->zmask.p, c->zw.p, c->kr.p, c->cap.p,
                                                     JvArgs(), nullptr);
    if (c->dim == 3)
        elem_kernel<3, 2><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, c->conn.p, c->ezone.p, c->kmul.p,
                                                       c->emask.p, c->eorient.p, c->xyz.p, psi, c->kr.p, nullptr,
                                                       c->qface.p);
    else
        elem_kernel<2, 2><<<blocks(T, 256), 256, 0, s>>>(T, N, c->zt, c->rho0, c->conn.p, c->ezone.p, c->kmul.p,
                                                       c->emask.p, c->eorient.p, c->xyz.p, psi, c->kr.p, nullptr,
                                                       c->qface.p);
    edge_flux_kernel<<<blocks(c->NE, 256), 256, 0, s>>>(c->NE, c->slot_ptr.p, c->slot.p, c->qface.p, c->ea.p,
                                                        c->eflux.p, c->wf.p);
    if (c->dim == 3)
        mask_kernel<3><<<blocks(c->T, 256), 256, 0, s>>>(c->T, c->eedge.p, c->eorient.p, c->wf.p, c->emask.p);
    else
        mask_kernel<2><<<blocks(c->T, 256), 256, 0, s>>>(c->T, c->eedge.p, c->eorient.p, c->wf.p, c->emask.p);
    FVP_HIP(hipGetLastError());
    c->averaging = 0;
    return 0;
}

extern "C" int fvp_get_upwind(fvp_ctx *c, double *w_front, double *edge_flux) {
    if (!c) { set_error("null context"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    FVP_HIP(hipStreamSynchronize(c->stream));
    if (w_front) FVP_HIP(hipMemcpy(w_front, c->wf.p, sizeof(double) * c->NE, hipMemcpyDeviceToHost));
    if (edge_flux) FVP_HIP(hipMemcpy(edge_flux, c->eflux.p, sizeof(double) * c->NE, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fvp_set_upwind(fvp_ctx *c, const double *w_front) {
    if (!c || !w_front) { set_error("null argument"); return -1; }
    int halves = 0;
    for (int i = 0; i < c->NE; ++i) {
        double w = w_front[i];
        if (w == 0.5) ++halves;
        else if (w != 0.0 && w != 1.0) { set_error("edge weights must be 0, 0.5 or 1"); return -1; }
    }
    if (halves && halves != c->NE) { set_error("averaged and upwind edge weights cannot be mixed"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    FVP_HIP(hipStreamSynchronize(c->stream));
    FVP_HIP(hipMemcpy(c->wf.p, w_front, sizeof(double) * c->NE, hipMemcpyHostToDevice));
    if (halves) { c->averaging = 1; return 0; }
    if (c->dim == 3)
        mask_kernel<3><<<blocks(c->T, 256), 256, 0, c->stream>>>(c->T, c->eedge.p, c->eorient.p, c->wf.p, c->emask.p);
    else
        mask_kernel<2><<<blocks(c->T, 256), 256, 0, c->stream>>>(c->T, c->eedge.p, c->eorient.p, c->wf.p, c->emask.p);
    FVP_HIP(hipGetLastError());
    c->averaging = 0;
    return 0;
}

extern "C" int fvp_capacity(fvp_ctx *c, double *ahh) {
    if (!c || !ahh) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    FVP_HIP(hipMemcpyAsync(ahh, c->cap.p, sizeof(double) * c->N, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

extern "C" int fvp_profile(fvp_ctx *c, int enable) {
    if (!c) { set_error("null context"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    FVP_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < kProfKinds; ++k) { c->prof_used[k] = 0; c->prof_ms[k] = 0.; c->prof_n[k] = 0; }
    c->prof = enable ? 1 : 0;
    return 0;
}

extern "C" int fvp_profile_get(fvp_ctx *c, int kind, double *total_ms, long *count) {
    if (!c || kind < 0 || kind >= kProfKinds) { set_error("invalid argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    prof_fold(c, kind);
    if (total_ms) *total_ms = c->prof_ms[kind];
    if (count) *count = c->prof_n[kind];
    return 0;
}

extern "C" int fvp_set_residual_launch(fvp_ctx *c, int variant, int tile_nodes, int xcd_map) {
    if (!c) { set_error("null context"); return -1; }
    if (variant != FVP_RESIDUAL_OWNER && variant != FVP_RESIDUAL_WINDOW) { set_error("unknown residual variant"); return -1; }
    if (tile_nodes < 64 || tile_nodes > 4096 || (tile_nodes & (tile_nodes - 1))) {
        set_error("tile_nodes must be a power of two in [64, 4096]");
        return -1;
    }
    FVP_HIP(hipSetDevice(c->device));
    FVP_HIP(hipStreamSynchronize(c->stream));
    const int old = c->tile_nodes, old_v = c->rvariant;
    c->rvariant = variant;  // the window variant's LDS check depends on it
    if ((tile_nodes != old || variant != old_v) && build_tiles(c, tile_nodes)) {
        const std::string why = fvp_last_error();
        c->rvariant = old_v;
        if (build_tiles(c, old)) return -1;  // keep the context usable
        set_error(why);
        return -1;
    }
    c->xcd_map = xcd_map ? 1 : 0;
    return 0;
}

// Launch-shape autotuner for the residual: every (tile size, XCD grouping)
// candidate that fits the LDS is built and timed with HIP events over a few
// upwinded residual evaluations at the given state; the fastest is kept.
extern "C" int fvp_autotune(fvp_ctx *c, double t, const double *psi, const double *psip, int *variant,
                            int *tile_nodes, int *xcd_map, double *best_us) {
    if (!c || !psi || !psip) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (fvp_begin_step(c, t, psi, psip)) return -1;
    hipEvent_t e0, e1;
    FVP_HIP(hipEventCreate(&e0));
    FVP_HIP(hipEventCreate(&e1));
    const int tiles[] = {256, 512, 1024, 2048};
    int best_t = c->tile_nodes, best_x = c->xcd_map, best_v = c->rvariant;
    double best = 1e300;
    const int reps = 10;
    for (int v : {FVP_RESIDUAL_OWNER, FVP_RESIDUAL_WINDOW})
    for (int tn : tiles) {
        if (fvp_set_residual_launch(c, v, tn, 1)) continue;  // window does not fit the LDS
        for (int x = 0; x <= 1; ++x) {
            c->xcd_map = x;
            for (int w = 0; w < 3; ++w)
                if (residual_impl(c, t, psi, psip, c->tmp3.p, 0)) return -1;
            FVP_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r)
                if (residual_impl(c, t, psi, psip, c->tmp3.p, 0)) return -1;
            FVP_HIP(hipEventRecord(e1, s));
            FVP_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FVP_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / reps;
            if (us < best) { best = us; best_t = tn; best_x = x; best_v = v; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (fvp_set_residual_launch(c, best_v, best_t, best_x)) return -1;
    if (variant) *variant = best_v;
    if (tile_nodes) *tile_nodes = best_t;
    if (xcd_map) *xcd_map = best_x;
    if (best_us) *best_us = best;
    return 0;
}
