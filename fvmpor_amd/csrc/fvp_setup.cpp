// Context construction: mesh ingest and the per-node / per-element tables the
// residual kernels read.  This replaces the host-side set-up of the reference
// (mesh::Mesh construction, mesh.cpp:70-80, 369-703, 786-1154;
// VarSatPhysicsImpl::initialise_vectors / initialise_shape_functions,
// fvmpor.h:346-644, 947-1124) with compact arrays: CV-face geometry is not
// stored, the kernels rebuild it from node coordinates on every call.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "fvp_internal.h"

namespace fvp {

static std::mutex g_err_mu;
static std::string g_err;

void set_error(const std::string &msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
}

template <class T>
int DevBuf<T>::alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return 0;
    hipError_t e = hipMalloc(&p, sizeof(T) * count);
    if (e != hipSuccess) {
        p = nullptr;
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}
template <class T>
int DevBuf<T>::upload(const T *h, size_t count) {
    if (alloc(count)) return -1;
    if (count == 0) return 0;
    hipError_t e = hipMemcpy(p, h, sizeof(T) * count, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}
template <class T>
void DevBuf<T>::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}
template struct DevBuf<int>;
template struct DevBuf<int4>;
template struct DevBuf<uint8_t>;
template struct DevBuf<double>;
template struct DevBuf<unsigned short>;
template struct DevBuf<float>;
template struct DevBuf<long>;
template struct DevBuf<unsigned>;

static const int TRI_EDGE[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int TET_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static const int TRI_FACE[3][3] = {{0, 1, -1}, {1, 2, -1}, {2, 0, -1}};
static const int TET_FACE[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

int upload_bc_values(fvp_ctx *c) {
    // rho0 * sum over the node's boundary CV faces of q.n A (fvmpor.h:721-753)
    std::vector<double> val(c->nb, 0.0);
    for (int k = 0; k < c->nb; ++k) {
        double s = 0.0;
        for (int q = c->bent_ptr[k]; q < c->bent_ptr[k + 1]; ++q) {
            const int tag = c->bent_tag[q];
            const double *b = &c->bc[(size_t)tag * FVP_BC_FIELDS];
            const double a = c->bent_area[q];
            if (a == 0.0) continue;
            switch (c->bc_type[tag]) {
            case 3:
            case 7: s += b[0] * a; break;
            case 6: {
                const double *nv = &c->bent_nrm[3 * (size_t)q];
                s += nv[0] * b[3] + nv[1] * b[4] + nv[2] * b[5];
                break;
            }
            default: break;
            }
        }
        val[k] = c->rho0 * s;
    }
    return c->bval.upload(val.data(), val.size());
}

// Dirichlet values per Dirichlet row from the current boundary rows: psi_D
// (type 1) or the hydrostatic head (1 + eta)(b - elevation) (type 4)
// (fvmpor.h:430-451, definitions.h:114-146).
int upload_dirichlet_values(fvp_ctx *c) {
    std::vector<double> dv(c->ndir);
    for (int k = 0; k < c->ndir; ++k) {
        const int tag = c->dtag_h[k];
        const double *b = &c->bc[(size_t)tag * FVP_BC_FIELDS];
        dv[k] = c->bc_type[tag] == 1 ? b[0] : (1. + b[2]) * (b[1] - c->delev_h[k]);
    }
    return c->dval.upload(dv.data(), dv.size());
}

// Node tiles: tile t owns nodes [t*TILE, (t+1)*TILE); it processes its primary
// elements (lowest node inside, a contiguous range) and the halo elements that
// reach into it from other tiles.  Rebuilt when the tile size changes.
int build_tiles(fvp_ctx *c, int TILE) {
    const int N = c->N, T = c->T, npe = c->npe;
    std::vector<int4> conn4(T);
    std::vector<uint8_t> ez(T);
    FVP_HIP(hipMemcpy(conn4.data(), c->conn.p, sizeof(int4) * T, hipMemcpyDeviceToHost));
    FVP_HIP(hipMemcpy(ez.data(), c->ezone.p, T, hipMemcpyDeviceToHost));
    {
        const int nt = (N + TILE - 1) / TILE;
        c->ntiles = nt;
        std::vector<int> eptr(nt + 1, 0), hcnt(nt + 1, 0);
        std::vector<int> first_tile(T);
        for (int e = 0; e < T; ++e) {
            const int4 q = conn4[e];
            int n[4] = {q.x, q.y, q.z, q.w};
            int m = n[0];
            for (int k = 1; k < npe; ++k) m = std::min(m, n[k]);
            first_tile[e] = m / TILE;
            eptr[m / TILE + 1]++;
        }
        for (int t = 0; t < nt; ++t) eptr[t + 1] += eptr[t];
        const bool window = c->rvariant == FVP_RESIDUAL_WINDOW;
        std::vector<std::vector<int>> halo(window ? nt : 0);
        for (int e = 0; window && e < T; ++e) {
            const int4 q = conn4[e];
            int n[4] = {q.x, q.y, q.z, q.w};
            int tl[4], cnt = 0;
            for (int k = 0; k < npe; ++k) {
                int t = n[k] / TILE;
                if (t == first_tile[e]) continue;
                bool seen = false;
                for (int r = 0; r < cnt; ++r) seen |= tl[r] == t;
                if (!seen) tl[cnt++] = t;
            }
            for (int r = 0; r < cnt; ++r) halo[tl[r]].push_back(e);
        }
        std::vector<int> hptr(nt + 1, 0), hlist;
        for (int t = 0; window && t < nt; ++t) hptr[t + 1] = hptr[t] + (int)halo[t].size();
        hlist.reserve(hptr[nt]);
        for (int t = 0; window && t < nt; ++t) hlist.insert(hlist.end(), halo[t].begin(), halo[t].end());
        // window of each tile = owned nodes [lo, hi) then its halo nodes (ascending);
        // per element visit (primary elements, then halo elements) the window-local
        // node indices, and the zones the tile's elements use (at most 4 per tile)
        std::vector<int> wptr(nt + 1, 0), whalo, vptr(nt + 1, 0), tz((size_t)nt * 4, -1);
        std::vector<unsigned short> lconn;
        lconn.reserve((size_t)4 * (T + hptr[nt]));
        std::vector<int> mark(N, -1), loc(N, 0);
        int wmax = 0, ntzmax = 1;
        for (int t = 0; window && t < nt; ++t) {
            const int lo = t * TILE, hi = std::min(N, lo + TILE), nown = hi - lo;
            std::vector<int> hal;
            unsigned zones = 0;
            auto scan = [&](int e) {
                const int4 q = conn4[e];
                int n[4] = {q.x, q.y, q.z, q.w};
                for (int k = 0; k < npe; ++k)
                    if ((n[k] < lo || n[k] >= hi) && mark[n[k]] != t) { mark[n[k]] = t; hal.push_back(n[k]); }
                zones |= 1u << ez[e];
            };
            for (int e = eptr[t]; e < eptr[t + 1]; ++e) scan(e);
            for (int e : halo[t]) scan(e);
            std::sort(hal.begin(), hal.end());
            for (size_t j = 0; j < hal.size(); ++j) loc[hal[j]] = nown + (int)j;
            const int win = nown + (int)hal.size();
            if (win > 65535) { set_error("tile window exceeds 65535 nodes"); return -1; }
            wmax = std::max(wmax, win);
            int nzt = 0;
            for (int z = 0; z < c->nz; ++z)
                if (zones & (1u << z)) {
                    if (nzt == 4) { set_error("a node tile touches more than 4 zones"); return -1; }
                    tz[(size_t)t * 4 + nzt++] = z;
                }
            ntzmax = std::max(ntzmax, nzt);
            auto emit = [&](int e) {
                const int4 q = conn4[e];
                int n[4] = {q.x, q.y, q.z, q.w};
                for (int k = 0; k < 4; ++k) {
                    int l = 0;
                    if (k < npe) l = (n[k] >= lo && n[k] < hi) ? n[k] - lo : loc[n[k]];
                    lconn.push_back((unsigned short)l);
                }
            };
            for (int e = eptr[t]; e < eptr[t + 1]; ++e) emit(e);
            for (int e : halo[t]) emit(e);
            whalo.insert(whalo.end(), hal.begin(), hal.end());
            wptr[t + 1] = (int)whalo.size();
            vptr[t + 1] = (int)(lconn.size() / 4);
        }
        c->win_max = wmax;
        c->ntz_max = ntzmax;
        c->tile_lds = window ? (size_t)(4 + ntzmax) * wmax * sizeof(double) + (size_t)TILE * sizeof(double)
                             : (size_t)TILE * sizeof(double);
        if (window) {
            if (c->t_wptr.upload(wptr.data(), nt + 1) || c->t_whalo.upload(whalo.data(), whalo.size()) ||
                c->t_vptr.upload(vptr.data(), nt + 1) || c->t_lconn.upload(lconn.data(), lconn.size()) ||
                c->t_tz.upload(tz.data(), tz.size()))
                return -1;
        } else {  // the element-owner variant reads none of the window tables
            c->t_wptr.release(); c->t_whalo.release(); c->t_vptr.release(); c->t_lconn.release(); c->t_tz.release();
        }
        std::vector<int> bptr(nt + 1, 0), dptr(nt + 1, 0);
        for (int k = 0; k < c->nb; ++k) bptr[c->bnode_h[k] / TILE + 1]++;
        std::vector<int> dn_h(c->ndir);
        if (c->ndir) FVP_HIP(hipMemcpy(dn_h.data(), c->dnode.p, sizeof(int) * c->ndir, hipMemcpyDeviceToHost));
        for (int k = 0; k < c->ndir; ++k) dptr[dn_h[k] / TILE + 1]++;
        for (int t = 0; t < nt; ++t) { bptr[t + 1] += bptr[t]; dptr[t + 1] += dptr[t]; }
        c->tile_elems = (long)T + hptr[nt];
        if (c->t_eptr.upload(eptr.data(), nt + 1) || c->t_hptr.upload(hptr.data(), nt + 1) ||
            (window && c->t_hlist.upload(hlist.data(), hlist.size())) || c->t_bptr.upload(bptr.data(), nt + 1) ||
            c->t_dptr.upload(dptr.data(), nt + 1))
            return -1;
        if (!window) c->t_hlist.release();
        // element-owner variant: contributions of a tile's elements to nodes of
        // other tiles are stored in export slots (numbered per tile, element order,
        // local node order) and gathered by their nodes in the finalisation
        if (!window) {
            std::vector<unsigned short> xoff(T);
            std::vector<int> xptr(nt + 1, 0), target;
            for (int t = 0; t < nt; ++t) {
                const int lo = t * TILE, hi = std::min(N, lo + TILE);
                int run = 0;
                for (int e = eptr[t]; e < eptr[t + 1]; ++e) {
                    if (run > 65535) { set_error("a residual tile exports more than 65535 contributions"); return -1; }
                    xoff[e] = (unsigned short)run;
                    const int4 q = conn4[e];
                    const int n[4] = {q.x, q.y, q.z, q.w};
                    for (int k = 0; k < npe; ++k)
                        if (n[k] < lo || n[k] >= hi) { target.push_back(n[k]); ++run; }
                }
                xptr[t + 1] = xptr[t] + run;
            }
            // slots are stored grouped by target node (exporting element order within
            // a node), so the finalisation reads each node's imports contiguously
            std::vector<int> iptr(N + 1, 0), xdst(target.size()), xtgt(target.size());
            for (int tn : target) iptr[tn + 1]++;
            for (int i = 0; i < N; ++i) iptr[i + 1] += iptr[i];
            std::vector<int> fill(iptr.begin(), iptr.end() - 1);
            for (size_t k = 0; k < target.size(); ++k) {
                xdst[k] = fill[target[k]]++;
                xtgt[xdst[k]] = target[k];
            }
            c->nexport = (long)target.size();
            if (c->t_xoff.upload(xoff.data(), T) || c->t_xptr.upload(xptr.data(), nt + 1) ||
                c->x_iptr.upload(iptr.data(), N + 1) || c->x_ilist.upload(xdst.data(), xdst.size()) ||
                c->x_tgt.upload(xtgt.data(), xtgt.size()) || c->xf.alloc(std::max<size_t>(1, target.size())))
                return -1;
        } else {
            c->t_xoff.release(); c->t_xptr.release(); c->x_iptr.release(); c->x_ilist.release(); c->xf.release();
            c->x_tgt.release();
            c->nexport = 0;
        }
    }
    c->tile_nodes = TILE;
    return tile_kernel_attrs(c);
}

}  // namespace fvp

using namespace fvp;

extern "C" const char *fvp_last_error(void) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

static int validate(const fvp_problem *p) {
    if (!p) { set_error("null problem"); return -1; }
    if (p->dim != 2 && p->dim != 3) { set_error("dim must be 2 or 3"); return -1; }
    if (p->n_nodes <= 0 || p->n_elements <= 0) { set_error("empty mesh"); return -1; }
    if (!p->xyz || !p->conn || !p->face_tags || !p->elem_zone || !p->zones) {
        set_error("missing mesh arrays"); return -1;
    }
    if (p->n_zones <= 0 || p->n_zones > kMaxZones) { set_error("n_zones must be 1..16"); return -1; }
    if (p->n_tags < 0 || p->n_tags > kMaxTags) { set_error("n_tags must be 0..65536"); return -1; }
    int npe = p->dim + 1;
    for (long i = 0; i < (long)p->n_elements * npe; ++i)
        if (p->conn[i] < 0 || p->conn[i] >= p->n_nodes) { set_error("connectivity out of range"); return -1; }
    for (int e = 0; e < p->n_elements; ++e)
        if (p->elem_zone[e] < 0 || p->elem_zone[e] >= p->n_zones) { set_error("element zone out of range"); return -1; }
    for (long i = 0; i < (long)p->n_elements * npe; ++i)
        if (p->face_tags[i] < 0 || (p->face_tags[i] > 0 && p->face_tags[i] >= p->n_tags)) {
            set_error("face tag without a boundary condition"); return -1;
        }
    return 0;
}

extern "C" int fvp_create(const fvp_problem *p, int device, fvp_ctx **out) {
    if (!out) { set_error("null output"); return -1; }
    *out = nullptr;
    if (validate(p)) return -1;
    FVP_HIP(hipSetDevice(device));
    fvp_ctx *c = new fvp_ctx();
    c->device = device;
    c->dim = p->dim; c->N = p->n_nodes; c->T = p->n_elements;
    c->npe = p->dim + 1; c->ned = p->dim == 2 ? 3 : 6;
    c->nz = p->n_zones; c->ntags = p->n_tags; c->rho0 = p->rho0;
    const int N = c->N, T = c->T, npe = c->npe, ned = c->ned, dim = c->dim;
    auto fail = [&]() { fvp_destroy(c); return -1; };
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed"); return fail();
    }
    c->stream = c->own_stream;

    c->zt.nz = c->nz;
    for (int z = 0; z < c->nz; ++z) {
        const double *r = p->zones + z * FVP_ZONE_FIELDS;
        double *d = c->zt.p[z];
        d[ZK_X] = r[0]; d[ZK_Y] = r[1]; d[ZK_Z] = r[2]; d[Z_PHI] = r[3];
        d[Z_ALPHA] = r[4]; d[Z_N] = r[5]; d[Z_M] = 1. - 1. / r[5]; d[Z_SR] = r[6]; d[Z_SS] = r[7];
    }
    c->bc_type.assign(p->bc_type, p->bc_type + c->ntags);
    c->bc.assign(p->bc, p->bc + (size_t)c->ntags * FVP_BC_FIELDS);

    // ---- element order: sorted by lowest node (stable), so the elements a node
    // tile owns are contiguous; every per-element table below uses this order
    std::vector<int> eperm(T), einv(T);  // new -> old, old -> new
    {
        std::vector<int> emin(T);
        for (int e = 0; e < T; ++e) {
            const int *q = p->conn + (size_t)e * npe;
            int m = q[0];
            for (int k = 1; k < npe; ++k) m = std::min(m, q[k]);
            emin[e] = m;
        }
        std::iota(eperm.begin(), eperm.end(), 0);
        std::stable_sort(eperm.begin(), eperm.end(), [&](int a, int b) { return emin[a] < emin[b]; });
        for (int e = 0; e < T; ++e) einv[eperm[e]] = e;
    }
    std::vector<int4> conn4(T);
    std::vector<uint8_t> ez(T);
    std::vector<double> km(T, 1.0);
    for (int e = 0; e < T; ++e) {
        const int o = eperm[e];
        const int *q = p->conn + (size_t)o * npe;
        conn4[e] = make_int4(q[0], q[1], q[2], dim == 3 ? q[3] : -1);
        ez[e] = (uint8_t)p->elem_zone[o];
        if (p->elem_kmul) km[e] = p->elem_kmul[o];
    }
    if (c->conn.upload(conn4.data(), T) || c->ezone.upload(ez.data(), T) || c->kmul.upload(km.data(), T))
        return fail();
    if (c->xyz.upload(p->xyz, (size_t)N * 3)) return fail();

    // ---- volumes: |T|/(d+1) per SCV (exact for simplices), zone fractions
    std::vector<double> V(N, 0.0);
    std::vector<double> evol(T);
    for (int e = 0; e < T; ++e) {
        const int *q = p->conn + (size_t)e * npe;
        const double *a = p->xyz + 3 * q[0], *b = p->xyz + 3 * q[1], *cc = p->xyz + 3 * q[2];
        double vol;
        if (dim == 2) {
            vol = 0.5 * std::fabs((b[0] - a[0]) * (cc[1] - a[1]) - (b[1] - a[1]) * (cc[0] - a[0]));
        } else {
            const double *d = p->xyz + 3 * q[3];
            double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
            double vx = cc[0] - a[0], vy = cc[1] - a[1], vz = cc[2] - a[2];
            double wx = d[0] - a[0], wy = d[1] - a[1], wz = d[2] - a[2];
            vol = std::fabs(ux * (vy * wz - vz * wy) - uy * (vx * wz - vz * wx) + uz * (vx * wy - vy * wx)) / 6.0;
        }
        if (!(vol > 0.0)) { set_error("degenerate element " + std::to_string(e)); return fail(); }
        evol[e] = vol / npe;
        for (int k = 0; k < npe; ++k) V[q[k]] += evol[e];
    }
    std::vector<double> invV(N);
    for (int i = 0; i < N; ++i) {
        if (!(V[i] > 0.0)) { set_error("node " + std::to_string(i) + " is in no element"); return fail(); }
        invV[i] = 1.0 / V[i];
    }
    std::vector<double> zw((size_t)c->nz * N, 0.0);
    std::vector<uint8_t> zmask(N, 0);
    for (int e = 0; e < T; ++e) {
        const int *q = p->conn + (size_t)e * npe;
        int z = p->elem_zone[e];
        for (int k = 0; k < npe; ++k) {
            zw[(size_t)z * N + q[k]] += evol[e] / V[q[k]];
            zmask[q[k]] |= (uint8_t)(1u << (z & 7));
        }
    }
    if (c->nz > 8) { set_error("more than 8 zones need a wider zone mask"); return fail(); }
    if (c->invV.upload(invV.data(), N) || c->zw.upload(zw.data(), zw.size()) ||
        c->zmask.upload(zmask.data(), N))
        return fail();
    if (c->kr.alloc((size_t)c->nz * N) || c->cap.alloc(N) || c->acc.alloc(N)) return fail();
    {
        // node records x, y, z, head: the coordinates are written here once, the
        // head by every residual
        std::vector<double> rec(4 * (size_t)N, 0.0);
        for (int i = 0; i < N; ++i)
            for (int d = 0; d < 3; ++d) rec[4 * (size_t)i + d] = p->xyz[3 * (size_t)i + d];
        if (c->nrec.upload(rec.data(), rec.size())) return fail();
    }
    FVP_HIP(hipMemset(c->kr.p, 0, sizeof(double) * c->kr.n));

    // ---- edges: sort (min, max, insertion index); first insertion fixes front/back
    {
        const size_t ns = (size_t)T * ned;
        std::vector<uint64_t> key(ns);
        std::vector<uint32_t> idx(ns);
        const int(*tab)[2] = dim == 2 ? TRI_EDGE : TET_EDGE;
        for (int e = 0; e < T; ++e)
            for (int j = 0; j < ned; ++j) {
                int n0 = p->conn[(size_t)e * npe + tab[j][0]], n1 = p->conn[(size_t)e * npe + tab[j][1]];
                uint32_t lo = (uint32_t)std::min(n0, n1), hi = (uint32_t)std::max(n0, n1);
                key[(size_t)e * ned + j] = ((uint64_t)lo << 32) | hi;
                idx[(size_t)e * ned + j] = (uint32_t)((size_t)e * ned + j);
            }
        std::vector<uint32_t> order(ns);
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            return key[a] != key[b] ? key[a] < key[b] : a < b;
        });
        std::vector<int> eedge(ns), slot(ns), slot_ptr;
        std::vector<uint8_t> orient(T, 0);
        slot_ptr.reserve(ns / 4 + 2);
        int ne = -1;
        int front = -1;
        for (size_t r = 0; r < ns; ++r) {
            uint32_t s = order[r];
            if (r == 0 || key[s] != key[order[r - 1]]) {
                ++ne;
                slot_ptr.push_back((int)r);
                int e = (int)(s / ned), j = (int)(s % ned);
                front = p->conn[(size_t)e * npe + tab[j][0]];  // first inserter's first node
            }
            int e = (int)(s / ned), j = (int)(s % ned);
            eedge[s] = ne;
            slot[r] = (int)s;
            if (p->conn[(size_t)e * npe + tab[j][0]] == front) orient[e] |= (uint8_t)(1u << j);
        }
        slot_ptr.push_back((int)ns);
        c->NE = ne + 1;
        // move the per-element tables to the sorted element order; slots keep their
        // (edge, original element) order, the order the reference sums CV faces in
        std::vector<int> eedge_n(ns), slot_n(ns);
        std::vector<uint8_t> orient_n(T);
        for (int e = 0; e < T; ++e) {
            const int o = eperm[e];
            orient_n[e] = orient[o];
            for (int j = 0; j < ned; ++j) eedge_n[(size_t)e * ned + j] = eedge[(size_t)o * ned + j];
        }
        for (size_t r = 0; r < ns; ++r) {
            int o = slot[r] / ned, j = slot[r] % ned;
            slot_n[r] = einv[o] * ned + j;
        }
        if (c->eedge.upload(eedge_n.data(), ns) || c->slot.upload(slot_n.data(), ns) ||
            c->slot_ptr.upload(slot_ptr.data(), slot_ptr.size()) || c->eorient.upload(orient_n.data(), T))
            return fail();
    }
    if (c->emask.alloc(T) || c->qface.alloc((size_t)T * ned) || c->ea.alloc(c->NE) || c->eflux.alloc(c->NE) ||
        c->wf.alloc(c->NE))
        return fail();
    FVP_HIP(hipMemset(c->emask.p, 0, T));
    FVP_HIP(hipMemset(c->eflux.p, 0, sizeof(double) * c->NE));
    {
        std::vector<double> half(c->NE, 0.5);
        FVP_HIP(hipMemcpy(c->wf.p, half.data(), sizeof(double) * c->NE, hipMemcpyHostToDevice));
    }
    c->averaging = 1;

    // ---- boundary CV faces aggregated per (node, tag)
    {
        std::vector<int> slot_of(N, -1);
        std::unordered_map<long, int> ent_of;  // (slot, tag) -> entry
        std::vector<int> e_slot, e_tag;
        std::vector<double> e_area, e_nrm;
        const int(*ftab)[3] = dim == 2 ? TRI_FACE : TET_FACE;
        const int nfe = npe, nfn = dim;
        for (int e = 0; e < T; ++e)
            for (int f = 0; f < nfe; ++f) {
                int tag = p->face_tags[(size_t)e * npe + f];
                if (!tag) continue;
                int fn[3];
                for (int q = 0; q < nfn; ++q) fn[q] = p->conn[(size_t)e * npe + ftab[f][q]];
                // face area and outward unit normal
                double A, nx, ny, nzv;
                const double *a = p->xyz + 3 * fn[0], *b = p->xyz + 3 * fn[1];
                if (dim == 2) {
                    double dx = b[0] - a[0], dy = b[1] - a[1];
                    A = std::sqrt(dx * dx + dy * dy);
                    nx = dy / A; ny = -dx / A; nzv = 0;
                } else {
                    const double *cc = p->xyz + 3 * fn[2];
                    double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
                    double vx = cc[0] - a[0], vy = cc[1] - a[1], vz = cc[2] - a[2];
                    double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
                    double l = std::sqrt(cx * cx + cy * cy + cz * cz);
                    A = 0.5 * l; nx = cx / l; ny = cy / l; nzv = cz / l;
                }
                // orient outwards: away from the element centroid
                double ce[3] = {0, 0, 0}, cf[3] = {0, 0, 0};
                for (int k = 0; k < npe; ++k)
                    for (int d = 0; d < 3; ++d) ce[d] += p->xyz[3 * p->conn[(size_t)e * npe + k] + d] / npe;
                for (int k = 0; k < nfn; ++k)
                    for (int d = 0; d < 3; ++d) cf[d] += p->xyz[3 * fn[k] + d] / nfn;
                if (nx * (cf[0] - ce[0]) + ny * (cf[1] - ce[1]) + nzv * (cf[2] - ce[2]) < 0) {
                    nx = -nx; ny = -ny; nzv = -nzv;
                }
                // each face node owns 1/dim of the face: a half segment (2D) or a third (3D)
                double Ai = A / nfn;
                for (int k = 0; k < nfn; ++k) {
                    int i = fn[k];
                    if (slot_of[i] < 0) {
                        slot_of[i] = c->nb++;
                        c->bnode_h.push_back(i);
                    }
                    const long key = (long)slot_of[i] * c->ntags + tag;
                    auto it = ent_of.find(key);
                    int q;
                    if (it == ent_of.end()) {
                        q = (int)e_slot.size();
                        ent_of.emplace(key, q);
                        e_slot.push_back(slot_of[i]);
                        e_tag.push_back(tag);
                        e_area.push_back(0.0);
                        e_nrm.insert(e_nrm.end(), {0.0, 0.0, 0.0});
                    } else q = it->second;
                    e_area[q] += Ai;
                    e_nrm[3 * (size_t)q] += Ai * Ai * nx;
                    e_nrm[3 * (size_t)q + 1] += Ai * Ai * ny;
                    e_nrm[3 * (size_t)q + 2] += Ai * Ai * nzv;
                }
            }
        // boundary slots in ascending node order (per-tile ranges below)
        std::vector<int> ord(c->nb);
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return c->bnode_h[a] < c->bnode_h[b]; });
        std::vector<int> bn(c->nb), rank(c->nb);
        for (int k = 0; k < c->nb; ++k) {
            bn[k] = c->bnode_h[ord[k]];
            rank[ord[k]] = k;
        }
        c->bnode_h.swap(bn);
        // entries grouped by sorted slot, ascending tag within a slot
        std::vector<int> eo(e_slot.size());
        std::iota(eo.begin(), eo.end(), 0);
        std::sort(eo.begin(), eo.end(), [&](int a, int b) {
            const int ra = rank[e_slot[a]], rb = rank[e_slot[b]];
            return ra != rb ? ra < rb : e_tag[a] < e_tag[b];
        });
        c->bent_ptr.assign(c->nb + 1, 0);
        c->bent_tag.resize(eo.size());
        c->bent_area.resize(eo.size());
        c->bent_nrm.resize(3 * eo.size());
        for (size_t q = 0; q < eo.size(); ++q) {
            const int src = eo[q];
            c->bent_ptr[rank[e_slot[src]] + 1]++;
            c->bent_tag[q] = e_tag[src];
            c->bent_area[q] = e_area[src];
            for (int d = 0; d < 3; ++d) c->bent_nrm[3 * q + d] = e_nrm[3 * (size_t)src + d];
        }
        for (int k = 0; k < c->nb; ++k) c->bent_ptr[k + 1] += c->bent_ptr[k];
        if (c->bnode.upload(c->bnode_h.data(), c->nb)) return fail();
        if (upload_bc_values(c)) return fail();
    }

    // ---- Dirichlet rows (fvmpor.h:404-451)
    {
        std::vector<int> dn;
        c->dtag_h.clear();
        c->delev_h.clear();
        if (p->node_tag_ptr && p->node_tags) {
            for (int i = 0; i < N; ++i) {
                int dtag = 0;
                for (int k = p->node_tag_ptr[i]; k < p->node_tag_ptr[i + 1]; ++k) {
                    int tag = p->node_tags[k];
                    if (tag > 0 && tag < c->ntags && (c->bc_type[tag] == 1 || c->bc_type[tag] == 4)) dtag = tag;
                }
                if (!dtag) continue;
                dn.push_back(i);
                c->dtag_h.push_back(dtag);
                c->delev_h.push_back(dim == 2 ? p->xyz[3 * i + 1] : p->xyz[3 * i + 2]);
            }
        }
        c->ndir = (int)dn.size();
        if (c->dnode.upload(dn.data(), dn.size()) || upload_dirichlet_values(c)) return fail();
    }
    if (build_tiles(c, c->tile_nodes) || edge_areas(c)) return fail();
    if (c->tmp1.alloc(N) || c->tmp2.alloc(N) || c->tmp3.alloc(N) || c->red.alloc(4096)) return fail();
    c->red_h.resize(4096);
    *out = c;
    return 0;
}

extern "C" int fvp_destroy(fvp_ctx *c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    jac_free(c);
    prof_release(c);
    c->conn.release(); c->ezone.release(); c->emask.release(); c->eorient.release(); c->kmul.release();
    c->xyz.release(); c->invV.release(); c->zmask.release(); c->zw.release(); c->kr.release();
    c->cap.release(); c->nrec.release(); c->acc.release(); c->bnode.release(); c->bval.release(); c->dnode.release();
    c->dval.release(); c->t_wptr.release(); c->t_whalo.release(); c->t_vptr.release(); c->t_lconn.release();
    c->t_tz.release(); c->t_eptr.release(); c->t_hptr.release(); c->t_hlist.release(); c->t_bptr.release();
    c->t_dptr.release(); c->t_xoff.release(); c->t_xptr.release(); c->x_iptr.release(); c->x_ilist.release();
    c->x_tgt.release();
    c->xf.release(); c->eedge.release(); c->slot_ptr.release(); c->slot.release(); c->qface.release();
    c->ea.release(); c->eflux.release(); c->wf.release(); c->tmp1.release(); c->tmp2.release();
    c->tmp3.release(); c->red.release();
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return 0;
}

extern "C" int fvp_sizes(const fvp_ctx *c, int *s) {
    if (!c || !s) { set_error("null argument"); return -1; }
    s[0] = c->N; s[1] = c->T; s[2] = c->NE; s[3] = c->ndir; s[4] = c->dim; s[5] = c->nz;
    return 0;
}

extern "C" int fvp_set_stream(fvp_ctx *c, void *stream) {
    if (!c) { set_error("null context"); return -1; }
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    return 0;
}

extern "C" int fvp_synchronize(fvp_ctx *c) {
    if (!c) { set_error("null context"); return -1; }
    FVP_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int fvp_set_bc(fvp_ctx *c, const double *bc) {
    if (!c || !bc) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    FVP_HIP(hipStreamSynchronize(c->stream));
    c->bc.assign(bc, bc + (size_t)c->ntags * FVP_BC_FIELDS);
    return upload_bc_values(c) || upload_dirichlet_values(c) ? -1 : 0;
}

// ---------------------------------------------------------------------------
// Reverse Cuthill-McKee renumbering of the node graph of a simplex mesh
// (Mesh::reorder_nodes_edges and find_RCM_from_edges, mesh.cpp:118-209,
// 1571-1615, which call Boost's cuthill_mckee_ordering and reverse it).  Per
// connected component: a pseudo-peripheral start node (George and Liu: from
// the component's lowest-degree node, breadth-first level structures, each
// time restarting from the lowest-degree node of the last level while the
// eccentricity grows), then a breadth-first sweep that appends each node's
// unvisited neighbours in increasing degree (ties in adjacency order); the
// concatenated order is reversed.  Degrees count the self loop the reference
// adds to every node (mesh.cpp:127-128).
namespace {

void node_graph(int N, int T, int npe, const int *conn, std::vector<int> &ia, std::vector<int> &ja) {
    std::vector<std::pair<int, int>> pr;
    pr.reserve((size_t)T * npe * (npe - 1));
    for (int e = 0; e < T; ++e)
        for (int a = 0; a < npe; ++a)
            for (int b = 0; b < npe; ++b)
                if (a != b) pr.emplace_back(conn[(size_t)e * npe + a], conn[(size_t)e * npe + b]);
    std::sort(pr.begin(), pr.end());
    pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
    ia.assign(N + 1, 0);
    ja.resize(pr.size());
    for (size_t k = 0; k < pr.size(); ++k) {
        ia[pr[k].first + 1]++;
        ja[k] = pr[k].second;
    }
    for (int i = 0; i < N; ++i) ia[i + 1] += ia[i];
}

// breadth-first level structure from s over unvisited nodes: returns the
// eccentricity; last = the last level's nodes
int bfs_levels(int s, const std::vector<int> &ia, const std::vector<int> &ja, const std::vector<char> &done,
               std::vector<int> &stamp, int mark, std::vector<int> &last) {
    std::vector<int> cur{s}, next;
    stamp[s] = mark;
    int ecc = 0;
    for (;;) {
        next.clear();
        for (int v : cur)
            for (int k = ia[v]; k < ia[v + 1]; ++k) {
                const int w = ja[k];
                if (!done[w] && stamp[w] != mark) { stamp[w] = mark; next.push_back(w); }
            }
        if (next.empty()) break;
        cur.swap(next);
        ++ecc;
    }
    last = cur;
    return ecc;
}

}  // namespace

extern "C" int fvp_rcm(int n_nodes, int n_elements, int nodes_per_element, const int *conn, int *perm) {
    if (n_nodes < 0 || n_elements < 0 || nodes_per_element < 2 || nodes_per_element > 4 || !perm ||
        (n_elements > 0 && !conn)) {
        set_error("fvp_rcm: invalid arguments");
        return -1;
    }
    const int N = n_nodes;
    for (long k = 0; k < (long)n_elements * nodes_per_element; ++k)
        if (conn[k] < 0 || conn[k] >= N) { set_error("fvp_rcm: node index out of range"); return -1; }
    std::vector<int> ia, ja;
    node_graph(N, n_elements, nodes_per_element, conn, ia, ja);
    std::vector<int> deg(N);
    for (int i = 0; i < N; ++i) deg[i] = ia[i + 1] - ia[i] + 1;  // self loop
    std::vector<char> done(N, 0);
    std::vector<int> stamp(N, -1), order, last, nbr;
    order.reserve(N);
    // nodes by increasing degree (ties by index): the first unvisited one seeds a component
    std::vector<int> bydeg(N);
    std::iota(bydeg.begin(), bydeg.end(), 0);
    std::stable_sort(bydeg.begin(), bydeg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    int mark = 0;
    for (int seed : bydeg) {
        if (done[seed]) continue;
        // pseudo-peripheral node
        int s = seed;
        int ecc = bfs_levels(s, ia, ja, done, stamp, mark++, last);
        for (;;) {
            int best = last[0];
            for (int v : last)
                if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
            const std::vector<int> prev = last;
            const int e2 = bfs_levels(best, ia, ja, done, stamp, mark++, last);
            if (e2 <= ecc) { last = prev; break; }
            s = best;
            ecc = e2;
        }
        // Cuthill-McKee sweep of the component
        size_t head = order.size();
        order.push_back(s);
        done[s] = 1;
        while (head < order.size()) {
            const int v = order[head++];
            nbr.clear();
            for (int k = ia[v]; k < ia[v + 1]; ++k)
                if (!done[ja[k]]) { done[ja[k]] = 1; nbr.push_back(ja[k]); }
            std::stable_sort(nbr.begin(), nbr.end(), [&](int a, int b) { return deg[a] < deg[b]; });
            order.insert(order.end(), nbr.begin(), nbr.end());
        }
    }
    // reversed: new node i is old node perm[i]
    for (int i = 0; i < N; ++i) perm[i] = order[N - 1 - i];
    return 0;
}
