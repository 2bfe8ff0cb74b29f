// Sparse linear algebra and the Newton-Krylov inner loop on gfx950.
//
//  * CSR SpMV: 8 lanes per row, shuffle reduction (rows hold ~7 (2D) to ~15 (3D) entries).
//  * ILU(0): level-scheduled factorisation and triangular solves.  Levels come
//    from the dependency DAG of L (forward) and U (backward); rows of a level
//    are independent and run as one launch.
//  * Jacobi: inverse diagonal.
//  * GMRES: right-preconditioned, restarted (gmres.c:26-201) with classical
//    Gram-Schmidt fused into one multi-dot and one multi-axpy pass per sweep
//    (applied twice when re-orthogonalising), so each Arnoldi step reads the
//    basis twice instead of 2(j+1) times as modified Gram-Schmidt would.  The
//    (m+1) x m Hessenberg system is reduced with Givens rotations on the
//    device; the host reads each step's residual estimate one step late.
//  * Coloured finite-difference Jacobian (preconditioner_dss.cpp:284-372): one
//    residual per colour of a distance-2 colouring.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <array>
#include <cmath>
#include <complex>
#include <functional>
#include <mutex>
#include <numeric>
#include <set>
#include <utility>
#include <vector>

#include "fvp_blas.h"
#include "fvp_dense.h"
#include "fvp_internal.h"

using fvp::DevBuf;
using fvp::set_error;

struct fvp_schur;

struct fvp_csr {
    int device = 0;
    hipStream_t stream = nullptr, own_stream = nullptr;
    int n = 0, nnz = 0;
    std::vector<int> ia_h, ja_h;
    DevBuf<int> ia, ja;
    DevBuf<double> a;
    // preconditioner
    int ptype = -1;
    DevBuf<double> lu, jac;
    DevBuf<int> diag;
    bool levels = false;
    std::vector<int> lptr_h, uptr_h;
    DevBuf<int> lrows, urows;
    DevBuf<int> flag;  // device error flag (zero pivot)
    // block-Jacobi ILU(0): per-block local CSR (columns relative to the block start)
    bool bj = false;
    int bs = FVP_BLOCK_ROWS;                // owned rows per block
    std::vector<int> bstart;                // explicit block boundaries (nb + 1), empty: every bs rows
    int overlap = 0;                        // layers of neighbouring rows added to each block (RAS)
    int nblk = 0, b_maxnz = 0, b_maxrows = 0, lds_max = 0;
    int factor_mode = -1;                   // staging form of the last block factorisation (bj_factor)
    DevBuf<int> b_roff, b_rowid, b_own;     // per block: first local row; global row of each local row; owned range
    DevBuf<double> b_zn2;                   // per block: its share of ||z||^2 of the last scaled solve
    DevBuf<int> b_ia, b_src, b_diag;        // row pointers into the block arrays, source entry, diagonal position
    DevBuf<unsigned short> b_ja;            // local column
    DevBuf<double> b_lu;
    DevBuf<float> b_lu32;                   // FP32 factor (FVP_PRECON_BJILU0_F32)
    DevBuf<int> b_lev_l, b_lev_u;           // per block: level pointer range [b_lev[b], b_lev[b+1]) into *_lptr
    DevBuf<int> b_lptr_l, b_lptr_u;         // level -> rows range in *_rows
    DevBuf<unsigned short> b_rows_l, b_rows_u;  // local rows grouped by level
    DevBuf<unsigned> fac_rec, fac_pair;     // factorisation: per entry its pair record (2 words), the pairs
    DevBuf<long> fac_pbase;                 // per block: its first pair
    // the solve's levels: the factor's levels cut into pieces of at most 32 rows
    DevBuf<int> s_lev_l, s_lev_u, s_lptr_l, s_lptr_u;
    int s_levmax_l = 0, s_levmax_u = 0;
    int s_maxslots = 0;                     // most slots of one block's triangle
    // packed solve layout per triangle (see bj_solve): per slot its meta word
    // pair, per block its entry base, per entry its source in b_lu (-1:
    // padding) and local column, the values of the current factor; per U slot
    // the source of its diagonal and the inverted diagonal
    DevBuf<unsigned> meta_l, meta_u;
    DevBuf<long> ebase_l, ebase_u;
    DevBuf<int> esrc_l, esrc_u, dsrc;
    DevBuf<unsigned short> col_l, col_u;
    DevBuf<double> val_l, val_u, dinv;
    DevBuf<float> val32_l, val32_u, dinv32;
    // per block its overflow entries (row parts past kBjRow): packed position and triangle
    DevBuf<int> ov_ptr;
    DevBuf<long> ov_pos;
    DevBuf<uint8_t> ov_tri;
    int ov_max = 0;
    // workspace
    int m_alloc = -1;
    DevBuf<double> Q, w, r, z, part, scal, hdev;
    DevBuf<double> Hd, gd, csd, snd, yd;    // Hessenberg (column-major, m+1 rows), Givens pairs, rhs, solution
    double *hstat = nullptr;                // pinned host: per Arnoldi step {residual estimate, h_{j+1,j}}
    DevBuf<double> Qk;                      // deflated GMRES: the kept harmonic Ritz vectors
    std::vector<hipEvent_t> ev;             // per Arnoldi step: its status copy has landed
    fvp_schur *sc = nullptr;                // Schur-complement preconditioner (FVP_PRECON_SCHUR)
};

// One-level Schur-complement preconditioner over subdomains of `bs`
// consecutive rows.  Interior unknowns (all couplings inside their subdomain)
// are numbered first, subdomain by subdomain, then the boundary unknowns.
struct fvp_schur {
    int nI = 0, nG = 0, nsub = 0, maxg = 0;
    fvp_csr *B = nullptr;                   // interior block: block-Jacobi ILU(0), one block per subdomain
    fvp::DevBuf<int> irow, grow;            // original row of each interior / boundary unknown
    fvp::DevBuf<int> bsrc;                  // source entry in A of each B entry
    fvp::DevBuf<int> e_ia, e_ja, e_loc, e_src;  // E (interior x boundary): column = boundary unknown, its
                                            // index inside the subdomain, source entry in A
    fvp::DevBuf<int> f_ia, f_ja, f_src;     // F (boundary x interior): column = interior unknown
    fvp::DevBuf<double> ev, fv;             // current values of E and F
    fvp::DevBuf<int> c_src;                 // C_ii entries: source in A and position in the dense blocks
    fvp::DevBuf<long> c_pos;
    fvp::DevBuf<int> goff, gsub, glocal;    // first boundary unknown per subdomain; subdomain and local index
    fvp::DevBuf<long> soff;                 // per subdomain: offset of its dense block (column major)
    fvp::DevBuf<double> S;                  // S_i, inverted in place by the build
    fvp::DevBuf<double> t1, t2, yG;         // work vectors (nI, nI, nG)
};

namespace fvp {

constexpr int kThreads = 256;

// ---- kernels ------------------------------------------------------------------
template <int LANES>
__global__ void __launch_bounds__(256) csr_gemv_kernel(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                       const double *__restrict__ a, double alpha,
                                                       const double *__restrict__ x, double beta,
                                                       double *__restrict__ y) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int row = tid / LANES, lane = tid % LANES;
    double s = 0.;
    if (row < n)
        for (int k = ia[row] + lane; k < ia[row + 1]; k += LANES) s += a[k] * x[ja[k]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LANES);
    if (row < n && lane == 0) y[row] = beta == 0. ? alpha * s : alpha * s + beta * y[row];
}

__global__ void ilu0_factor_level(int cnt, const int *__restrict__ rows, const int *__restrict__ ia,
                                  const int *__restrict__ ja, const int *__restrict__ diag, double *__restrict__ lu,
                                  int *__restrict__ flag) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    int i = rows[t];
    int end = ia[i + 1];
    for (int kk = ia[i]; kk < diag[i]; ++kk) {
        int k = ja[kk];
        int kpos = ia[k], kend = ia[k + 1];
        double lik = lu[kk] * lu[diag[k]];
        lu[kk] = lik;
        for (int jj = kk + 1; jj < end; ++jj) {
            int j = ja[jj];
            while (kpos < kend && ja[kpos] < j) kpos++;
            if (kpos == kend) break;
            if (ja[kpos] == j) lu[jj] -= lik * lu[kpos];
        }
    }
    double d = lu[diag[i]];
    if (d == 0.) { atomicOr(flag, 1); lu[diag[i]] = 0.; }
    else lu[diag[i]] = 1. / d;
}

template <int LANES>
__global__ void __launch_bounds__(256) ilu0_lower_level(int cnt, const int *__restrict__ rows,
                                                        const int *__restrict__ ia, const int *__restrict__ ja,
                                                        const int *__restrict__ diag, const double *__restrict__ lu,
                                                        const double *__restrict__ x, double *__restrict__ y) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int t = tid / LANES, lane = tid % LANES;
    int i = t < cnt ? rows[t] : 0;
    double s = 0.;
    if (t < cnt)
        for (int k = ia[i] + lane; k < diag[i]; k += LANES) s += lu[k] * y[ja[k]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LANES);
    if (t < cnt && lane == 0) y[i] = x[i] - s;
}

template <int LANES>
__global__ void __launch_bounds__(256) ilu0_upper_level(int cnt, const int *__restrict__ rows,
                                                        const int *__restrict__ ia, const int *__restrict__ ja,
                                                        const int *__restrict__ diag, const double *__restrict__ lu,
                                                        double *__restrict__ y) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int t = tid / LANES, lane = tid % LANES;
    int i = t < cnt ? rows[t] : 0;
    double s = 0.;
    if (t < cnt)
        for (int k = diag[i] + 1 + lane; k < ia[i + 1]; k += LANES) s += lu[k] * y[ja[k]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LANES);
    if (t < cnt && lane == 0) y[i] = (y[i] - s) * lu[diag[i]];
}

__global__ void jacobi_build(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             const double *__restrict__ a, double *__restrict__ J, int *__restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int pos = ia[i];
    while (pos < ia[i + 1] && ja[pos] < i) pos++;
    if (pos == ia[i + 1] || ja[pos] != i || a[pos] == 0.) { atomicOr(flag, 1); J[i] = 0.; return; }
    J[i] = 1. / a[pos];
}

__global__ void vmul(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] * b[i];
}

__global__ void vdiv(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] / b[i];
}

__global__ void vcopy(long n, const double *__restrict__ a, double *__restrict__ b) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// y = b - y
__global__ void vrsub(long n, const double *__restrict__ b, double *__restrict__ y) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = b[i] - y[i];
}

// y += x
__global__ void vadd(long n, const double *__restrict__ x, double *__restrict__ y) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// dst = src / sqrt(nsq[0])
__global__ void vscale_inv_norm(long n, const double *__restrict__ src, const double *__restrict__ nsq,
                                double *__restrict__ dst) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] * (1. / sqrt(nsq[0]));
}

constexpr int kChunk = 16;  // basis columns per Gram-Schmidt dot pass

__device__ __forceinline__ double wave_reduce(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Workgroups of the Gram-Schmidt passes: 512 x 256 threads (2 per CU), each
// thread streaming several row pairs per column.  Measured on the 1M-node
// box's node graph (30 unpreconditioned GMRES steps): 4.62 ms with 2048
// workgroups, 4.53 with 1024, 4.42-4.47 with 512, 4.42 with 256; fewer
// workgroups also mean fewer partials to reduce.
constexpr int kGsBlocks = 512;

// Leading dimension of the basis: columns padded to an even length, so that
// every column starts 16-byte aligned for the pairwise loads.
static inline long gs_ldq(long n) { return (n + 1) & ~1L; }
static inline int gs_blocks(long n) {
    return (int)std::max(1L, std::min((long)kGsBlocks, ((n >> 1) + kThreads - 1) / kThreads));
}

__device__ __forceinline__ void block_partial(double v, double *__restrict__ out) {
    __shared__ double lds[kThreads / 64];
    v = wave_reduce(v);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int k = 0; k < kThreads / 64; ++k) t += lds[k];
        *out = t;
    }
}

// Block partials of K values at once: every wavefront reduces all K with
// shuffles, one barrier, then thread c sums the wavefronts' values of column c
// into part[c * gridDim.x + block].
template <int K>
__device__ __forceinline__ void block_partials(const double (&acc)[K], int nk, double *__restrict__ part) {
    __shared__ double lds[kThreads / 64][K];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const double v = wave_reduce(acc[c]);
        if (lane == 0) lds[wv][c] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nk) {
        double t = 0.;
        for (int k = 0; k < kThreads / 64; ++k) t += lds[k][threadIdx.x];
        part[threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// Block partials of w . Q[:, c] for c < NC and of w . w (stored as column NC
// when ww is set): part[c * gridDim.x + block].  Each thread reads element
// pairs (16-byte loads; the basis columns are padded to an even length, ldq),
// every column load of a pair issued before the first product.  Skipped
// entirely when skip[0] == 0 (second pass not needed).
template <int NC>
__global__ void __launch_bounds__(kThreads) mdot_kernel(long n, const double *__restrict__ w,
                                                        const double *__restrict__ Q, long ldq, int ww,
                                                        double *__restrict__ part, const double *__restrict__ skip) {
    if (skip && skip[0] == 0.) return;
    double acc[NC + 1];
#pragma unroll
    for (int c = 0; c <= NC; ++c) acc[c] = 0.;
    const long np = n >> 1;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
        const double2 a = reinterpret_cast<const double2 *>(w)[p];
        double2 q[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) q[c] = reinterpret_cast<const double2 *>(Q + c * ldq)[p];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a.x * q[c].x + a.y * q[c].y;
        acc[NC] += a.x * a.x + a.y * a.y;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const double a = w[n - 1];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a * Q[c * ldq + n - 1];
        acc[NC] += a * a;
    }
    block_partials<NC + 1>(acc, ww ? NC + 1 : NC, part);
}

// w -= Q[:, 0..NC) coef, and block partials of the updated w . w into
// part[block]; element pairs as in mdot_kernel.
template <int NC>
__global__ void __launch_bounds__(kThreads) axpy_norm(long n, double *__restrict__ w, const double *__restrict__ Q,
                                                      long ldq, const double *__restrict__ coef,
                                                      double *__restrict__ part, const double *__restrict__ skip) {
    if (skip && skip[0] == 0.) return;
    double h[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) h[c] = coef[c];
    double acc = 0.;
    const long np = n >> 1;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
        double2 q[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) q[c] = reinterpret_cast<const double2 *>(Q + c * ldq)[p];
        double2 a = reinterpret_cast<double2 *>(w)[p];
        double sx = 0., sy = 0.;
#pragma unroll
        for (int c = 0; c < NC; ++c) { sx += h[c] * q[c].x; sy += h[c] * q[c].y; }
        a.x -= sx;
        a.y -= sy;
        reinterpret_cast<double2 *>(w)[p] = a;
        acc += a.x * a.x + a.y * a.y;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        double sx = 0.;
#pragma unroll
        for (int c = 0; c < NC; ++c) sx += h[c] * Q[c * ldq + n - 1];
        const double a = w[n - 1] - sx;
        w[n - 1] = a;
        acc += a * a;
    }
    block_partial(acc, part + blockIdx.x);
}

// The first pass's update and the second pass's dots in one read of the basis:
// w -= Q[:, 0..NC) coef, then block partials of Q[:, c] . w (c < NC) and of
// w . w (column NC) on the updated w.
template <int NC>
__global__ void __launch_bounds__(kThreads) axpy_mdot(long n, double *__restrict__ w, const double *__restrict__ Q,
                                                      long ldq, const double *__restrict__ coef,
                                                      double *__restrict__ part) {
    double h[NC], acc[NC + 1];
#pragma unroll
    for (int c = 0; c < NC; ++c) h[c] = coef[c];
#pragma unroll
    for (int c = 0; c <= NC; ++c) acc[c] = 0.;
    const long np = n >> 1;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
        double2 q[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) q[c] = reinterpret_cast<const double2 *>(Q + c * ldq)[p];
        double2 a = reinterpret_cast<double2 *>(w)[p];
        double sx = 0., sy = 0.;
#pragma unroll
        for (int c = 0; c < NC; ++c) { sx += h[c] * q[c].x; sy += h[c] * q[c].y; }
        a.x -= sx;
        a.y -= sy;
        reinterpret_cast<double2 *>(w)[p] = a;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a.x * q[c].x + a.y * q[c].y;
        acc[NC] += a.x * a.x + a.y * a.y;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        double qt[NC], sx = 0.;
#pragma unroll
        for (int c = 0; c < NC; ++c) { qt[c] = Q[c * ldq + n - 1]; sx += h[c] * qt[c]; }
        const double a = w[n - 1] - sx;
        w[n - 1] = a;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a * qt[c];
        acc[NC] += a * a;
    }
    block_partials<NC + 1>(acc, NC + 1, part);
}

using MdotLaunch = void (*)(hipStream_t, int, long, const double *, const double *, long, int, double *,
                           const double *);
using AxpyLaunch = void (*)(hipStream_t, int, long, double *, const double *, long, const double *, double *,
                           const double *);
template <int NC>
static void mdot_launch(hipStream_t s, int nb, long n, const double *w, const double *Q, long ldq, int ww,
                        double *part, const double *skip) {
    mdot_kernel<NC><<<nb, kThreads, 0, s>>>(n, w, Q, ldq, ww, part, skip);
}
template <int NC>
static void axpy_launch(hipStream_t s, int nb, long n, double *w, const double *Q, long ldq, const double *coef,
                        double *part, const double *skip) {
    axpy_norm<NC><<<nb, kThreads, 0, s>>>(n, w, Q, ldq, coef, part, skip);
}
template <int NC>
static void axpy_mdot_launch(hipStream_t s, int nb, long n, double *w, const double *Q, long ldq, const double *coef,
                             double *part, const double *) {
    axpy_mdot<NC><<<nb, kThreads, 0, s>>>(n, w, Q, ldq, coef, part);
}
template <int... I>
static constexpr std::array<AxpyLaunch, sizeof...(I)> axpy_mdot_table(std::integer_sequence<int, I...>) {
    return {&axpy_mdot_launch<I + 1>...};
}
template <int... I>
static constexpr std::array<MdotLaunch, sizeof...(I)> mdot_table(std::integer_sequence<int, I...>) {
    return {&mdot_launch<I + 1>...};
}
template <int... I>
static constexpr std::array<AxpyLaunch, sizeof...(I)> axpy_table(std::integer_sequence<int, I...>) {
    return {&axpy_launch<I + 1>...};
}
// per column count 1..kChunk
static const auto kMdot = mdot_table(std::make_integer_sequence<int, kChunk>{});
static const auto kAxpy = axpy_table(std::make_integer_sequence<int, kChunk>{});
static const auto kAxpyMdot = axpy_mdot_table(std::make_integer_sequence<int, kChunk>{});

// Selective reorthogonalisation (Kahan-Parlett "twice is enough"): a second
// classical Gram-Schmidt pass when the first one cancelled more than half of
// ||w||^2 (mode 1), always (mode 2) or never (mode 0).
__global__ void reorth_decide(const double *__restrict__ ww, const double *__restrict__ n1, int mode,
                              double *__restrict__ flag) {
    if (threadIdx.x == 0) flag[0] = mode == 2 ? 1. : mode == 0 ? 0. : (n1[0] < 0.5 * ww[0] ? 1. : 0.);
}

// netlib drotg
__host__ __device__ inline void drotg_dev(double *da, double *db, double *c, double *s) {
    double roe = fabs(*da) > fabs(*db) ? *da : *db;
    double scale = fabs(*da) + fabs(*db), r, z;
    if (scale == 0.) { *c = 1.; *s = 0.; r = 0.; z = 0.; }
    else {
        double a = *da / scale, b = *db / scale;
        r = scale * sqrt(a * a + b * b);
        r = (roe >= 0 ? 1. : -1.) * r;
        *c = *da / r; *s = *db / r;
        z = 1.;
        if (fabs(*da) > fabs(*db)) z = *s;
        if (fabs(*db) >= fabs(*da) && *c != 0.) z = 1. / *c;
    }
    *da = r; *db = z;
}

// End of Arnoldi step j in one launch.  Every block forms 1/h_{j+1,j} from the
// same norm (||w||^2 after the first pass, or after the second when flag is
// set) and scales its slice of w into q_{j+1}.  Block 0 also sets the new
// Hessenberg column j = h1 (+ h2 after a second pass) with h_{j+1,j} = ||w||,
// applies the previous rotations and the new Givens rotation, rotates the
// right-hand side g, and writes {|g_{j+1}|, h_{j+1,j}} to stat and, when hstat
// is given (pinned host memory the device can write), there too.
// With npart, ||w||^2 is not reduced beforehand: every block sums the nb
// partials of the single Gram-Schmidt pass itself, in reduce_cols' order (the
// same value bit for bit, one launch fewer), and the slices are grid-strided.
// Workgroups of arnoldi_finish (grid-strided): few enough that every one
// summing the norm's partials costs little.
constexpr int kFinishBlocks = 512;
__global__ void __launch_bounds__(256) arnoldi_finish(int j, int m, const double *__restrict__ h1,
                                                      const double *__restrict__ h2, const double *__restrict__ n1,
                                                      const double *__restrict__ n2, const double *__restrict__ flag,
                                                      double *__restrict__ H, double *__restrict__ cs,
                                                      double *__restrict__ sn, double *__restrict__ g,
                                                      double *__restrict__ stat, double *__restrict__ hstat, long n,
                                                      const double *__restrict__ w, double *__restrict__ q,
                                                      const double *__restrict__ npart = nullptr, int nb = 0) {
    __shared__ double lds[4];
    const bool two = !npart && flag && flag[0] != 0.;
    double nn;
    if (npart) {
        // each lane's partials in order, four loads in flight at a time
        double s = 0.;
        const int bd = blockDim.x;
        for (int b0 = threadIdx.x; b0 < nb; b0 += 4 * bd) {
            double q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = b0 + u * bd < nb ? npart[b0 + u * bd] : 0.;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (b0 + u * bd < nb) s += q[u];
        }
        s = wave_reduce(s);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
        __syncthreads();
        double v = 0.;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) v += lds[k];
        nn = v;
    } else {
        nn = two ? n2[0] : n1[0];
    }
    const double hn = sqrt(nn);
    const double ninv = hn > 0. ? 1. / hn : 0.;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        q[i] = w[i] * ninv;
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double *hc = H + (size_t)j * (m + 1);
    for (int k = 0; k <= j; ++k) hc[k] = two ? h1[k] + h2[k] : h1[k];
    hc[j + 1] = hn;
    for (int k = 0; k < j; ++k) {
        const double t1 = cs[k] * hc[k] + sn[k] * hc[k + 1];
        hc[k + 1] = cs[k] * hc[k + 1] - sn[k] * hc[k];
        hc[k] = t1;
    }
    double a = hc[j], b = hc[j + 1], c, sv;
    drotg_dev(&a, &b, &c, &sv);
    cs[j] = c; sn[j] = sv;
    hc[j] = c * hc[j] + sv * hc[j + 1];
    hc[j + 1] = 0.;
    const double gj = g[j];
    g[j] = c * gj;
    g[j + 1] = -sv * gj;
    stat[0] = fabs(g[j + 1]);
    stat[1] = hn;
    if (hstat) {
        hstat[0] = stat[0];
        hstat[1] = hn;
        __threadfence_system();
    }
}

__global__ void gmres_cycle_init(double *__restrict__ g, int len, double beta) {
    for (int i = threadIdx.x; i < len; i += blockDim.x) g[i] = i == 0 ? beta : 0.;
}

// y = R^{-1} g on the first k rotated columns (back substitution)
__global__ void gmres_backsolve(int k, int m, const double *__restrict__ H, const double *__restrict__ g,
                                double *__restrict__ y) {
    if (threadIdx.x != 0) return;
    for (int i = k - 1; i >= 0; --i) {
        double sum = g[i];
        for (int l = i + 1; l < k; ++l) sum -= H[(size_t)l * (m + 1) + i] * y[l];
        y[i] = sum / H[(size_t)i * (m + 1) + i];
    }
}

// out[c] (+)= sum over blocks of part[c * nb + block]
// With flag set, the block of the last column (||w||^2 after a first
// Gram-Schmidt pass) also decides the second pass: flag = 1 when that norm
// fell below half of ww[0] (||w||^2 before the pass), or always when ww is null.
__global__ void __launch_bounds__(kThreads) reduce_cols(int nb, const double *__restrict__ part,
                                                        double *__restrict__ out, int accumulate,
                                                        const double *__restrict__ skip,
                                                        const double *__restrict__ ww = nullptr,
                                                        double *__restrict__ flag = nullptr) {
    if (skip && skip[0] == 0.) return;
    __shared__ double lds[kThreads / 64];
    int c = blockIdx.x;
    // each lane's partials in order, four loads in flight at a time
    double s = 0.;
    const int bd = blockDim.x;
    const double *pc = part + (size_t)c * nb;
    for (int b0 = threadIdx.x; b0 < nb; b0 += 4 * bd) {
        double q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = b0 + u * bd < nb ? pc[b0 + u * bd] : 0.;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (b0 + u * bd < nb) s += q[u];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) v += lds[k];
        out[c] = accumulate ? out[c] + v : v;
        if (flag && c == (int)gridDim.x - 1) flag[0] = (!ww || v < 0.5 * ww[0]) ? 1. : 0.;
    }
}

// w += sign * sum_c coef[c] Q[:, c]
// out = sum_c y[c] Q[:, c]  (y on the device)
__global__ void lincomb(long n, double *__restrict__ out, const double *__restrict__ Q, long ldq, int nc,
                        const double *__restrict__ y) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.;
    for (int c = 0; c < nc; ++c) s += Q[c * ldq + i] * y[c];
    out[i] = s;
}

// y = a x + b y  (host scalars)
__global__ void vaxpby(long n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = b == 0. ? a * x[i] : a * x[i] + b * y[i];
}

// ---- block-Jacobi ILU(0): one workgroup per block of bs consecutive rows
// (factorisation: bj_factor; the rows of one level run concurrently)
constexpr int kBjFactorThreads = 512;

// compiler-only ordering of one wavefront's LDS accesses (the LDS executes a
// wavefront's accesses in issue order; no wait for outstanding memory loads)
__device__ __forceinline__ void wave_lds_order() { asm volatile("" ::: "memory"); }

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The update pairs of the factorisation are precomputed from the pattern: for
// the L entry e = (i, k) of a block, rec[e] = {first pair, count | pos(u_kk) << 16}
// and each pair packs the block positions of u_kj (low half) and a_ij (high
// half) for the columns j > k that rows i and k share.  A group of kBjGroup
// lanes factors one row (IKJ order: for each L entry in column order,
// l_ik = a_ik u_kk^-1, then a_ij -= l_ik u_kj over the entry's pairs, one pair
// per lane); kBjFactorThreads / kBjGroup rows of a level are in flight per
// workgroup.  The records of a row are held in the group's registers and the
// pairs of the next L entry are loaded while the current one is applied.
// STAGED: the block's values are staged in LDS; otherwise the factorisation
// works in place on the block's values in memory.
constexpr int kBjGroup = 16;

__device__ __forceinline__ uint2 group_shfl(uint2 v, int src) {
    return make_uint2((unsigned)__shfl((int)v.x, src, kBjGroup), (unsigned)__shfl((int)v.y, src, kBjGroup));
}

template <typename T, bool STAGED>
__global__ void __launch_bounds__(kBjFactorThreads) bj_factor(
    const int *__restrict__ roff, const int *__restrict__ ia, const int *__restrict__ diag,
    const int *__restrict__ src, const double *__restrict__ a, const uint2 *__restrict__ rec,
    const long *__restrict__ pbase, const unsigned *__restrict__ pairs, const int *__restrict__ blev,
    const int *__restrict__ lptr, const unsigned short *__restrict__ rows, T *__restrict__ lu,
    int *__restrict__ flag) {
    extern __shared__ double smem[];
    const int b = blockIdx.x;
    const int base = roff[b];
    const int cnt = roff[b + 1] - base;
    const int z0 = ia[base], nzb = ia[base + cnt] - z0;
    T *val = STAGED ? (T *)smem : lu + z0;
    for (int k = threadIdx.x; k < nzb; k += blockDim.x) val[k] = (T)a[src[z0 + k]];
    if (!STAGED) __threadfence_block();
    __syncthreads();
    const int gl = threadIdx.x & (kBjGroup - 1), g = threadIdx.x / kBjGroup, ng = blockDim.x / kBjGroup;
    const unsigned *pr = pairs + pbase[b];
    const uint2 *rb = rec + z0;
    for (int l = blev[b]; l < blev[b + 1]; ++l) {
        for (int t = lptr[l] + g; t < lptr[l + 1]; t += ng) {
            const int r = rows[t];
            const int s0 = ia[base + r] - z0, d = diag[base + r] - z0 - s0;
            uint2 rc = gl < d ? rb[s0 + gl] : make_uint2(0u, 0u);  // records kk = gl (mod kBjGroup)
            uint2 cur = group_shfl(rc, 0);
            unsigned q = gl < (cur.y & 0xffffu) ? pr[cur.x + gl] : 0u;
            for (int kk = 0; kk < d; ++kk) {
                const int k1 = kk + 1;
                if (k1 < d && (k1 & (kBjGroup - 1)) == 0) rc = k1 + gl < d ? rb[s0 + k1 + gl] : make_uint2(0u, 0u);
                uint2 nxt = make_uint2(0u, 0u);
                unsigned qn = 0u;
                if (k1 < d) {
                    nxt = group_shfl(rc, k1 & (kBjGroup - 1));
                    qn = gl < (nxt.y & 0xffffu) ? pr[nxt.x + gl] : 0u;
                }
                const int np = (int)(cur.y & 0xffffu);
                const T lik = val[s0 + kk] * val[cur.y >> 16];  // u_kk is stored inverted
                if (gl < np) val[q >> 16] -= lik * val[q & 0xffffu];
                for (int j = gl + kBjGroup; j < np; j += kBjGroup) {
                    const unsigned q2 = pr[cur.x + j];
                    val[q2 >> 16] -= lik * val[q2 & 0xffffu];
                }
                wave_lds_sync();
                if (gl == 0) val[s0 + kk] = lik;
                wave_lds_sync();
                cur = nxt;
                q = qn;
            }
            if (gl == 0) {
                const T piv = val[s0 + d];
                if (piv == T(0)) { atomicOr(flag, 1); val[s0 + d] = T(0); }
                else val[s0 + d] = T(1) / piv;
            }
            wave_lds_sync();
        }
        if (!STAGED) __threadfence_block();
        __syncthreads();
    }
    if (STAGED)
        for (int k = threadIdx.x; k < nzb; k += blockDim.x) lu[z0 + k] = val[k];
}

// Solve: one wavefront per block.  Only the block's solution, its overflow
// entries and its level pointers live in LDS (so that every block of the
// matrix is resident at once and the CU interleaves their dependency chains).
// The factor streams from a level-ordered packed copy per triangle: the row
// parts one after another (each padded to an even length), so the bytes read
// are the factor's own.  Per row part (a "slot", level order) a meta word pair
// gives its start and its overflow offset, its row and its length.  A lane
// pair takes a slot: the pair reads the fixed window of kBjRow entries
// starting at the slot's start, one half each (past the row's length the
// window runs into the following rows, which the level uses anyway, and is
// masked); the halves' sums meet by a swap within the pair, and the even lane
// subtracts the entries past kBjRow (from LDS) and stores the row.  So a level
// of up to 32 slots costs one half window of instructions per lane.  T is the
// factor's arithmetic type (double, or float for the FP32 preconditioner); x
// and y stay FP64.
constexpr int kBjRow = 12;
#ifndef BJ_AHEAD_F32
#define BJ_AHEAD_F32 8
#endif
#ifndef BJ_AHEAD_F64
#define BJ_AHEAD_F64 4
#endif

// After a factorisation: the packed values (src < 0: padding, zero) and, for
// U, every slot's inverted diagonal.
template <typename T>
__global__ void bj_pack_vals(long n, const int *__restrict__ src, const T *__restrict__ lu, T *__restrict__ val) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int q = src[k];
    val[k] = q >= 0 ? lu[q] : T(0);
}

constexpr int kBjHalf = kBjRow / 2;  // window entries per lane of a pair

template <typename T>
struct BjWin {
    T a[kBjHalf];
    unsigned c[kBjHalf / 2];  // local columns, two per word
    unsigned m0, m1;          // meta words
    T d;                      // inverted diagonal (U)
    bool have;
};

// x plus the other lane of the pair (lanes 2i, 2i+1): a DPP quad permutation
// [1, 0, 3, 2], so both lanes get the same sum
__device__ __forceinline__ float pair_sum(float x) {
    const int y = __builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false);
    return x + __int_as_float(y);
}
__device__ __forceinline__ double pair_sum(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return x + __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
struct BjMeta {
    unsigned m0, m1;  // start (low 16) | overflow offset (high 16); row (low 16) | length (high 16)
    int s;            // slot
    bool have;
};

template <typename T, bool UPPER>
__device__ __forceinline__ void bj_win_load(BjWin<T> &w, const BjMeta &m, int half, const T *__restrict__ val,
                                            const unsigned short *__restrict__ col, const T *__restrict__ dinv) {
    // this lane's half of the window: even start, so pairs stay aligned
    const long k = (long)(m.m0 & 0xffffu) + half * kBjHalf;
    if (sizeof(T) == 4) {
        const uint2 *p = reinterpret_cast<const uint2 *>(val + k);
#pragma unroll
        for (int i = 0; i < kBjHalf / 2; ++i) {
            const uint2 q = p[i];
            w.a[2 * i] = (T)__uint_as_float(q.x);
            w.a[2 * i + 1] = (T)__uint_as_float(q.y);
        }
    } else {
        const double2 *p = reinterpret_cast<const double2 *>(val + k);
#pragma unroll
        for (int i = 0; i < kBjHalf / 2; ++i) {
            const double2 q = p[i];
            w.a[2 * i] = (T)q.x;
            w.a[2 * i + 1] = (T)q.y;
        }
    }
    const unsigned *pc = reinterpret_cast<const unsigned *>(col + k);
#pragma unroll
    for (int i = 0; i < kBjHalf / 2; ++i) w.c[i] = pc[i];
    w.m0 = m.m0;
    w.m1 = m.m1;
    w.d = UPPER ? dinv[m.s] : T(0);
    w.have = m.have;
}

// Each lane of the pair sums its half window (bj_win_partial); the even lane
// then finishes the row (bj_win_finish): s = v[r] - (both halves) - (entries
// past kBjRow, in column order), times the inverted diagonal for U.  Every
// lane takes part in the pair sum (lanes without a slot read a valid clamped
// slot and store nothing).
// The half window's partial sum; vr: v[r], read with the window's columns
// (rows of one level do not depend on each other, so v[r] is final here).
template <typename T>
__device__ __forceinline__ T bj_win_partial(const BjWin<T> &w, int half, const T *v, T &vr) {
    const int r = (int)(w.m1 & 0xffffu), kn = (int)(w.m1 >> 16);
    T p = T(0);
#pragma unroll
    for (int u = 0; u < kBjHalf; ++u) {
        // past the row's length the window holds the following rows' entries
        // (possibly of another block): masked to 0 * v[r]
        const int ug = half * kBjHalf + u;
        const unsigned cw = w.c[u / 2];
        const int c = ug < kn ? (int)((u & 1) ? (cw >> 16) : (cw & 0xffffu)) : r;
        const T a = ug < kn ? w.a[u] : T(0);
        p += a * v[c];
    }
    vr = v[r];
    return p;
}

// OVG: the overflow entries are read from the packed arrays in memory (blocks
// whose overflow lists do not fit the LDS) instead of from LDS
template <typename T, bool UPPER, bool OVG>
__device__ __forceinline__ void bj_win_finish(const BjWin<T> &w, int half, T p, T vr, T *v, T *vd,
                                              const uint4 *ovp, const T *__restrict__ val,
                                              const unsigned short *__restrict__ col) {
    const int r = (int)(w.m1 & 0xffffu), kn = (int)(w.m1 >> 16);
    // every lane forms and stores s (lanes without a row store to their own
    // spare slot vd), so v[r]'s read stays with the window's reads
    const bool mine = w.have && !half;
    T s = vr - pair_sum(p);
    if (mine && kn > kBjRow) {
        if (OVG) {  // overflow entries straight after the window in the packed arrays
            const int k0 = (int)(w.m0 & 0xffffu);
            for (int k = k0 + kBjRow; k < k0 + kn; ++k) s -= val[k] * v[col[k]];
        } else {
            const int oo = (int)(w.m0 >> 16);
            for (int k = oo; k < oo + kn - kBjRow; ++k) {
                const uint4 q = ovp[k];
                T a;
                if (sizeof(T) == 4) a = (T)__uint_as_float(q.x);
                else a = (T)__longlong_as_double((long long)(((unsigned long long)q.y << 32) | q.x));
                s -= a * v[q.z];
            }
        }
    }
    if (UPPER) s *= w.d;  // inverted diagonal
    *(mine ? v + r : vd) = s;
}

// levels loaded ahead: deep enough to cover a memory round trip with the short
// per-level arithmetic; FP64 windows are larger, so fewer of them fit the registers
template <typename T>
constexpr int bj_ahead() { return sizeof(T) == 4 ? BJ_AHEAD_F32 : BJ_AHEAD_F64; }

// Staging a triangle's meta words (its ns slots from meta) in LDS: one round
// is kBjMeta loads per lane, all in flight together, so a block of up to
// 64 * kBjMeta slots per triangle waits out one memory latency.
constexpr int kBjMeta = 32;
__device__ __forceinline__ void bj_meta_load(uint2 (&q)[kBjMeta], const uint2 *__restrict__ meta, int ns, int k0) {
#pragma unroll
    for (int i = 0; i < kBjMeta; ++i) q[i] = meta[min(k0 + 64 * i, ns - 1)];
}
__device__ __forceinline__ void bj_meta_store(const uint2 (&q)[kBjMeta], uint2 *ms, int ns, int k0) {
#pragma unroll
    for (int i = 0; i < kBjMeta; ++i)
        if (k0 + 64 * i < ns) ms[k0 + 64 * i] = q[i];
}
// the slots from k0 (this lane's first) on, round by round
__device__ __forceinline__ void bj_meta_stage(const uint2 *__restrict__ meta, uint2 *ms, int ns, int k0) {
    for (; k0 < ns; k0 += 64 * kBjMeta) {
        uint2 q[kBjMeta];
        bj_meta_load(q, meta, ns, k0);
        bj_meta_store(q, ms, ns, k0);
    }
}

// One triangle of one block: lane pair t/2 takes the (t/2)-th slot of each
// level (the solve's levels hold at most 32 slots).  The triangle's meta words
// are staged in LDS (ms, slot lp[0] first) before the sweep, so a level's
// window addresses come from LDS and only the windows themselves are loads in
// flight: kA levels ahead, one window per lane and level, each consumed by the
// same unrolled stage that issued it kA stages earlier (no loop-carried
// register holds a pending load, so the compiler's wait counts keep all kA
// levels in flight).  No global load with a data-dependent count is issued
// inside the level loop, and every pass runs all kA levels (those past the
// last have no lane with work).
template <typename T, bool UPPER, bool OVG>
__device__ __forceinline__ void bj_sweep(T *v, T *vd, int nl, const int *lp, const uint2 *ms,
                                         const T *__restrict__ val, const unsigned short *__restrict__ col,
                                         const T *__restrict__ dinv, const uint4 *ovp) {
    const int t = threadIdx.x;
    if (nl <= 0) return;
    constexpr int kA = bj_ahead<T>();
    const int s0 = lp[0];
    const int q = t >> 1, half = t & 1;
    // a level's slot range, then this lane's slot and its meta words
    auto level_range = [&](int l, int &p0, int &cnt) {
        const int lc = min(l, nl - 1);
        p0 = lp[lc];
        cnt = lp[lc + 1] - p0;
    };
    auto slot_meta = [&](int l, int p0, int cnt, BjMeta &m) {
        m.s = p0 + min(q, cnt - 1);
        const uint2 mq = ms[m.s - s0];
        m.m0 = mq.x;
        m.m1 = mq.y;
        m.have = l < nl && q < cnt;
    };
    BjWin<T> wr[kA];
#pragma unroll
    for (int d = 0; d < kA; ++d) {
        int p0, cnt;
        level_range(d, p0, cnt);
        BjMeta m;
        slot_meta(d, p0, cnt, m);
        bj_win_load<T, UPPER>(wr[d], m, half, val, col, dinv);
        wave_lds_order();
    }
    // Per stage three LDS round trips overlap: level lb + d + kA's range is
    // read with the update's values, its meta words while the update's sums
    // are formed; the window of that level is issued last.
    for (int lb = 0; lb < nl; lb += kA) {
#pragma unroll
        for (int d = 0; d < kA; ++d) {
            const int ln = lb + d + kA;
            int p0, cnt;
            level_range(ln, p0, cnt);
            T vr;
            const T p = bj_win_partial<T>(wr[d], half, v, vr);
            BjMeta m;
            slot_meta(ln, p0, cnt, m);
            bj_win_finish<T, UPPER, OVG>(wr[d], half, p, vr, v, vd, ovp, val, col);
            bj_win_load<T, UPPER>(wr[d], m, half, val, col, dinv);  // level ln
            // one wavefront: the LDS executes its accesses in order, so only
            // the compiler must keep them in program order
            wave_lds_order();
        }
    }
}

// LDS bytes of one block's solve: the solution (T x maxrows) and a spare
// slot per lane (T x 64), the block's
// overflow entries (16 bytes each; none with OVG), its L and U level pointers,
// the meta words of one triangle's slots
__host__ __device__ inline size_t bj_solve_lds(size_t tsize, int maxrows, int ovmax, int levl, int levu,
                                                int maxslots) {
    return ((tsize * (size_t)(maxrows + 64) + 15) & ~(size_t)15) + 16 * (size_t)ovmax +
           ((4 * (size_t)(levl + levu + 2) + 7) & ~(size_t)7) + 8 * (size_t)maxslots;
}

// rowid: global row of each local row (nullptr: block b's rows are the
// contiguous range starting at roff[b]); own: the block's owned local rows,
// the only ones written back (restricted additive Schwarz).  One wavefront per
// block, so that every block of the matrix is resident at once.  (Staging the
// whole block in LDS was measured slower: it limits the resident blocks, and
// one wavefront stages too slowly.)  Per triangle: solve levels blev (per
// block) into lptr (per level, into the level-ordered slots, at most 32 per
// level), per slot its meta word pair, the packed values and local columns
// from the block's entry base ebase, U's inverted diagonals; the block's
// overflow entries are the packed entries at ov_pos[ov_ptr[b] ..).
template <typename T, bool OVG>
__global__ void __launch_bounds__(64) bj_solve(const int *__restrict__ roff, const int *__restrict__ rowid,
                                               const int *__restrict__ own, const int *__restrict__ blev_l,
                                               const int *__restrict__ lptr_l, const uint2 *__restrict__ meta_l,
                                               const long *__restrict__ ebase_l, const T *__restrict__ val_l,
                                               const unsigned short *__restrict__ col_l,
                                               const int *__restrict__ blev_u, const int *__restrict__ lptr_u,
                                               const uint2 *__restrict__ meta_u, const long *__restrict__ ebase_u,
                                               const T *__restrict__ val_u, const unsigned short *__restrict__ col_u,
                                               const T *__restrict__ dinv, const int *__restrict__ ov_ptr,
                                               const long *__restrict__ ov_pos, const unsigned char *__restrict__ ov_tri,
                                               int maxrows, int ovmax, int levmax_l, int levmax_u,
                                               const double *__restrict__ x,
                                               const double *__restrict__ xdiv, double *__restrict__ y,
                                               double *__restrict__ yn2) {
    extern __shared__ double smem[];
    const int b = blockIdx.x;
    const int base = roff[b];
    const int cnt = roff[b + 1] - base;
    char *lds = reinterpret_cast<char *>(smem);
    T *v = reinterpret_cast<T *>(lds);
    // v: the block's rows, then one spare slot per lane
    T *vd = v + maxrows + threadIdx.x;
    uint4 *ovp = reinterpret_cast<uint4 *>(lds + ((sizeof(T) * (size_t)(maxrows + 64) + 15) & ~(size_t)15));
    const int l0 = blev_l[b], nl = blev_l[b + 1] - l0;
    const int u0 = blev_u[b], nu = blev_u[b + 1] - u0;
    // L's meta words: the first round's loads are issued here, ahead of
    // staging x, and stored to the LDS after it
    const int sl0 = nl > 0 ? lptr_l[l0] : 0, nsl = nl > 0 ? lptr_l[l0 + nl] - sl0 : 0;
    uint2 mq[kBjMeta];
    if (nsl > 0) bj_meta_load(mq, meta_l + sl0, nsl, threadIdx.x);
    // stage x (optionally scaled by 1 / xdiv) in batches of kBjStage rows per
    // lane, so that a batch's gathers are in flight together
    constexpr int kBjStage = 16;
    for (int r0 = threadIdx.x; r0 < cnt; r0 += 64 * kBjStage) {
        int g[kBjStage];
        double xv[kBjStage], dv[kBjStage];
#pragma unroll
        for (int k = 0; k < kBjStage; ++k) {
            const int r = min(r0 + 64 * k, cnt - 1);
            g[k] = rowid ? rowid[base + r] : base + r;
        }
#pragma unroll
        for (int k = 0; k < kBjStage; ++k) {
            xv[k] = x[g[k]];
            dv[k] = xdiv ? xdiv[g[k]] : 1.;
        }
#pragma unroll
        for (int k = 0; k < kBjStage; ++k) {
            const int r = r0 + 64 * k;
            if (r < cnt) v[r] = (T)(xdiv ? xv[k] / dv[k] : xv[k]);
        }
    }
    // overflow entries: value bits, local column
    const int ob = ov_ptr[b], on = OVG ? 0 : ov_ptr[b + 1] - ob;
    if (OVG) ovmax = 0;
    for (int k = threadIdx.x; k < on; k += 64) {
        const long e = ov_pos[ob + k];
        const bool up = ov_tri[ob + k];
        const T a = up ? val_u[e] : val_l[e];
        unsigned lo, hi = 0u;
        if (sizeof(T) == 4) lo = __float_as_uint((float)a);
        else {
            const unsigned long long u = (unsigned long long)__double_as_longlong((double)a);
            lo = (unsigned)u;
            hi = (unsigned)(u >> 32);
        }
        ovp[k] = make_uint4(lo, hi, (unsigned)(up ? col_u[e] : col_l[e]), 0u);
    }
    int *lpl = reinterpret_cast<int *>(ovp + ovmax);
    int *lpu = lpl + nl + 1;
    uint2 *ms = reinterpret_cast<uint2 *>(reinterpret_cast<char *>(lpl) +
                                          ((4 * (size_t)(levmax_l + levmax_u + 2) + 7) & ~(size_t)7));
    for (int k = threadIdx.x; k <= nl; k += 64) lpl[k] = lptr_l[l0 + k];
    for (int k = threadIdx.x; k <= nu; k += 64) lpu[k] = lptr_u[u0 + k];
    if (nsl > 0) bj_meta_store(mq, ms, nsl, threadIdx.x);
    bj_meta_stage(meta_l + sl0, ms, nsl, threadIdx.x + 64 * kBjMeta);
    wave_lds_order();
    bj_sweep<T, false, OVG>(v, vd, nl, lpl, ms, val_l + ebase_l[b], col_l + ebase_l[b], nullptr, ovp);
    // U's meta words replace L's (the sweep's LDS accesses execute in order;
    // loading them during L's sweep instead was measured no faster)
    if (nu > 0) bj_meta_stage(meta_u + lpu[0], ms, lpu[nu] - lpu[0], threadIdx.x);
    wave_lds_order();
    bj_sweep<T, true, OVG>(v, vd, nu, lpu, ms, val_u + ebase_u[b], col_u + ebase_u[b], dinv, ovp);
    const int olo = own[2 * b], ohi = own[2 * b + 1];
    double q = 0.;
    for (int r0 = olo + threadIdx.x; r0 < ohi; r0 += 64 * kBjStage) {
        int g[kBjStage];
#pragma unroll
        for (int k = 0; k < kBjStage; ++k) {
            const int r = min(r0 + 64 * k, ohi - 1);
            g[k] = rowid ? rowid[base + r] : base + r;
        }
#pragma unroll
        for (int k = 0; k < kBjStage; ++k) {
            const int r = r0 + 64 * k;
            if (r < ohi) {
                const double yr = (double)v[r];
                y[g[k]] = yr;
                q += yr * yr;
            }
        }
    }
    if (yn2) {  // this block's share of ||y||^2 (owned rows partition the rows)
        q = wave_reduce(q);
        if (threadIdx.x == 0) yn2[b] = q;
    }
}

static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// ---- Schur-complement preconditioner (schur.c:887-981 form, :1583-1626 apply)
// Largest boundary set of one subdomain: its dense block is inverted in LDS.
constexpr int kSchurMax = 128;

__global__ void vgather(long n, const int *__restrict__ src, const double *__restrict__ in, double *__restrict__ out) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[src[i]];
}

__global__ void vscatter(long n, const int *__restrict__ dst, const double *__restrict__ in, double *__restrict__ out) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[dst[i]] = in[i];
}

// column c of every subdomain's E at once (schur_form_local's dense E column)
__global__ void schur_col_rhs(int nI, const int *__restrict__ e_ia, const int *__restrict__ e_loc,
                              const double *__restrict__ ev, int c, double *__restrict__ t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nI) return;
    double v = 0.;
    for (int k = e_ia[i]; k < e_ia[i + 1]; ++k)
        if (e_loc[k] == c) v = ev[k];
    t[i] = v;
}

// S_b(:, c) = -F_b y for every subdomain with more than c boundary unknowns
__global__ void schur_col_store(int nG, const int *__restrict__ f_ia, const int *__restrict__ f_ja,
                                const double *__restrict__ fv, const double *__restrict__ y,
                                const int *__restrict__ gsub, const int *__restrict__ glocal,
                                const int *__restrict__ goff, const long *__restrict__ soff, int c,
                                double *__restrict__ S) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nG) return;
    const int b = gsub[j], n = goff[b + 1] - goff[b];
    if (c >= n) return;
    double s = 0.;
    for (int k = f_ia[j]; k < f_ia[j + 1]; ++k) s += fv[k] * y[f_ja[k]];
    S[soff[b] + (long)c * n + glocal[j]] = -s;
}

__global__ void schur_add_c(long nc, const long *__restrict__ pos, const int *__restrict__ src,
                            const double *__restrict__ a, double *__restrict__ S) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k < nc) S[pos[k]] += a[src[k]];
}

// In-place Gauss-Jordan inversion of each subdomain's dense S_b (column major,
// n <= kSchurMax) in LDS, with partial pivoting: at step k the first row of
// largest |a(r, k)|, r >= k, is swapped into row k; the interchanges are undone
// as column swaps in reverse order at the end.  Waves own columns, lanes own
// rows (two rows per lane), so every LDS access of a wave is one contiguous
// column segment.  A zero pivot sets *flag.
__global__ void __launch_bounds__(256) schur_invert(const int *__restrict__ goff, const long *__restrict__ soff,
                                                    double *__restrict__ S, int *__restrict__ flag) {
    extern __shared__ double sm[];  // n x n block, pivot row, pivot column, n pivots, 2 flags
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int n = goff[b + 1] - goff[b];
    if (n == 0) return;
    double *a = sm, *rk = sm + n * n, *fk = rk + n;
    int *perm = reinterpret_cast<int *>(fk + n);
    int &s_piv = perm[n], &s_bad = perm[n + 1];
    double *g = S + soff[b];
    for (int e = tid; e < n * n; e += 256) a[e] = g[e];
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (wv == 0) {
            double best = -1.;
            int bi = n;
            for (int r = k + lane; r < n; r += 64) {
                const double v = fabs(a[k * n + r]);
                if (v > best) { best = v; bi = r; }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (lane == 0) {
                s_piv = bi;
                perm[k] = bi;
                if (!(best > 0.)) s_bad = 1;
            }
        }
        __syncthreads();
        if (s_bad) break;
        const int p = s_piv;
        if (p != k)
            for (int c = tid; c < n; c += 256) {
                const double t = a[c * n + k];
                a[c * n + k] = a[c * n + p];
                a[c * n + p] = t;
            }
        __syncthreads();
        const double inv = 1.0 / a[k * n + k];
        for (int c = tid; c < n; c += 256) {
            rk[c] = c == k ? inv : a[c * n + k] * inv;
            fk[c] = a[k * n + c];
        }
        __syncthreads();
        const double f0 = lane < n ? fk[lane] : 0., f1 = lane + 64 < n ? fk[lane + 64] : 0.;
        for (int c = wv; c < n; c += 4) {
            const double rc = rk[c];
            double *col = a + c * n;
            if (lane < n)
                col[lane] = lane == k ? rc : (c == k ? -f0 * inv : col[lane] - f0 * rc);
            if (lane + 64 < n)
                col[lane + 64] = lane + 64 == k ? rc : (c == k ? -f1 * inv : col[lane + 64] - f1 * rc);
        }
        __syncthreads();
    }
    if (s_bad) {
        if (tid == 0) atomicOr(flag, 1);
        return;
    }
    for (int k = n - 1; k >= 0; --k) {
        const int p = perm[k];
        if (p != k)
            for (int r = tid; r < n; r += 256) {
                const double t = a[k * n + r];
                a[k * n + r] = a[p * n + r];
                a[p * n + r] = t;
            }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += 256) g[e] = a[e];
}

// Middle of the apply, one workgroup per subdomain: v = x_G - F u (u = M_B^-1 x_I),
// y_G = S_b^-1 v, written to y (original rows) and to yG (boundary numbering).
__global__ void __launch_bounds__(kSchurMax) schur_mid(const int *__restrict__ goff, const long *__restrict__ soff,
                                                       const double *__restrict__ Sinv, const int *__restrict__ grow,
                                                       const int *__restrict__ f_ia, const int *__restrict__ f_ja,
                                                       const double *__restrict__ fv, const double *__restrict__ x,
                                                       const double *__restrict__ u, double *__restrict__ y,
                                                       double *__restrict__ yG) {
    __shared__ double v[kSchurMax];
    const int b = blockIdx.x, r = threadIdx.x;
    const int g0 = goff[b], n = goff[b + 1] - g0;
    if (r < n) {
        const int j = g0 + r;
        double s = x[grow[j]];
        for (int k = f_ia[j]; k < f_ia[j + 1]; ++k) s -= fv[k] * u[f_ja[k]];
        v[r] = s;
    }
    __syncthreads();
    if (r < n) {
        const double *Sb = Sinv + soff[b];
        double s = 0.;
        for (int c = 0; c < n; ++c) s += Sb[(long)c * n + r] * v[c];
        yG[g0 + r] = s;
        y[grow[g0 + r]] = s;
    }
}

// t_I = x_I - E y_G
__global__ void schur_back(int nI, const int *__restrict__ irow, const int *__restrict__ e_ia,
                           const int *__restrict__ e_ja, const double *__restrict__ ev, const double *__restrict__ x,
                           const double *__restrict__ yG, double *__restrict__ t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nI) return;
    double s = x[irow[i]];
    for (int k = e_ia[i]; k < e_ia[i + 1]; ++k) s -= ev[k] * yG[e_ja[k]];
    t[i] = s;
}

// ---- host helpers ---------------------------------------------------------------
static int ensure_work(fvp_csr *A, int m) {
    if (A->m_alloc >= m) return 0;
    long n = A->n;
    if (A->Q.alloc((size_t)gs_ldq(n) * (m + 1)) || A->w.alloc(n) || A->r.alloc(n) || A->z.alloc(n) ||
        A->part.alloc((size_t)kGsBlocks * (kChunk + 1)) || A->scal.alloc(16) || A->hdev.alloc(2 * (m + 2) + 2) ||
        A->Hd.alloc((size_t)(m + 1) * m) || A->gd.alloc(m + 1) || A->csd.alloc(m) || A->snd.alloc(m) ||
        A->yd.alloc(m))
        return -1;
    if (A->hstat) (void)hipHostFree(A->hstat);
    A->hstat = nullptr;
    // per Arnoldi step {estimate, h_{j+1,j}}, then two scalar slots
    if (hipHostMalloc((void **)&A->hstat, sizeof(double) * (2 * (m + 1) + 2), hipHostMallocDefault) != hipSuccess) {
        A->hstat = nullptr;
        set_error("pinned host allocation failed");
        return -1;
    }
    while ((int)A->ev.size() < m + 1) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { set_error("event creation failed"); return -1; }
        A->ev.push_back(e);
    }
    A->m_alloc = m;
    return 0;
}

// out[0..nc) = Q[:, 0..nc)^T w and out[nc] = w . w (device), chunks of kChunk
// columns; the w . w term rides on the last chunk.  skip: see mdot_kernel.
static void multidot(hipStream_t s, long n, const double *w, const double *Q, long ldq, int nc, double *part,
                     double *out, const double *skip) {
    const int nb = gs_blocks(n);
    for (int c0 = 0; c0 < nc; c0 += kChunk) {
        const int k = std::min(kChunk, nc - c0);
        const int last = c0 + k == nc;
        kMdot[k - 1](s, nb, n, w, Q + c0 * ldq, ldq, last, part, skip);
        reduce_cols<<<k + last, kThreads, 0, s>>>(nb, part, out + c0, 0, skip);
    }
}

// w -= Q[:, 0..nc) coef in chunks of kChunk columns; *norm2 = the updated w . w
// (norm2 null: left as the gs_blocks(n) partials in part)
static void multiaxpy(hipStream_t s, long n, double *w, const double *Q, long ldq, int nc, const double *coef,
                      double *part, double *norm2, const double *skip) {
    const int nb = gs_blocks(n);
    for (int c0 = 0; c0 < nc; c0 += kChunk) {
        const int k = std::min(kChunk, nc - c0);
        kAxpy[k - 1](s, nb, n, w, Q + c0 * ldq, ldq, coef + c0, part, skip);
    }
    if (norm2) reduce_cols<<<1, kThreads, 0, s>>>(nb, part, norm2, 0, skip);
}

using Op = std::function<int(const double *, double *)>;

// Restarted right-preconditioned GMRES (gmres.c:26-201) with the host kept off
// the critical path: the Hessenberg column, the Givens rotations and the
// residual estimate are updated on the device; each Arnoldi step's estimate is
// copied to pinned host memory and read one step later, while the next step
// is already queued (a cycle that has converged runs one extra, discarded
// step).  Classical Gram-Schmidt with a second pass only when the first one
// cancelled more than half of ||w||^2 (reorthogonalise = 1), always (2) or
// never (0).  A cycle ends early once the estimate meets tol * ||b||.
// Device and host memory of one GMRES solve: the basis Q (gs_ldq(n) x (m+1)),
// work vectors w, r, z (gs_ldq(n) each), reduction partials part
// (kGsBlocks x (kChunk+1)), scalars sc (16), per-step coefficients hdev
// (2 (m+2) + 2), the Hessenberg matrix Hd ((m+1) x m), Givens pairs, rhs and
// solution; hstat (host, 2 (m+1)) receives each step's {estimate, h_{j+1,j}}.
// With ev (one event per step) the host reads step j-1 while step j runs;
// without, each step's pair is read right after the step (same control flow
// and results, one synchronisation per step).
struct GmresWork {
    double *Q, *w, *r, *z, *part, *sc, *hdev, *Hd, *gd, *csd, *snd, *yd;
    double *hstat;
    const hipEvent_t *ev;
    bool hmapped;  // hstat is pinned host memory the kernels write directly
    // with hmapped: two pinned scalar slots for ||r||^2 and ||b||^2, and the
    // event the host polls for them (hipMemcpyAsync and a blocking wait otherwise)
    double *hscal;
    hipEvent_t hev;
};

// doubles of a GmresWork carved from one device buffer (each piece 256-byte aligned)
static size_t gmres_work_doubles(long n, int m) {
    auto up = [](size_t v) { return (v + 31) & ~(size_t)31; };
    const size_t ldq = (size_t)gs_ldq(n);
    return up(ldq * (m + 1)) + 3 * up(ldq) + up((size_t)kGsBlocks * (kChunk + 1)) + up(16) + up(2 * (m + 2) + 2) +
           up((size_t)(m + 1) * m) + 4 * up(m + 1);
}
static GmresWork gmres_work_carve(double *base, long n, int m) {
    auto up = [](size_t v) { return (v + 31) & ~(size_t)31; };
    const size_t ldq = (size_t)gs_ldq(n);
    GmresWork wk{};
    double *q = base;
    auto take = [&](size_t cnt) { double *r = q; q += up(cnt); return r; };
    wk.Q = take(ldq * (m + 1));
    wk.w = take(ldq);
    wk.r = take(ldq);
    wk.z = take(ldq);
    wk.part = take((size_t)kGsBlocks * (kChunk + 1));
    wk.sc = take(16);
    wk.hdev = take(2 * (m + 2) + 2);
    wk.Hd = take((size_t)(m + 1) * m);
    wk.gd = take(m + 1);
    wk.csd = take(m + 1);
    wk.snd = take(m + 1);
    wk.yd = take(m + 1);
    return wk;
}

static int gmres_core(const GmresWork &W, hipStream_t s, long n, const Op &A, const Op &M, const double *b,
                      double *x, const fvp_gmres_params *p, fvp_gmres_stats *st, bool x_zero = false);

int gmres_device(fvp_csr *W, hipStream_t s, long n, const Op &A, const Op &M, const double *b, double *x,
                 const fvp_gmres_params *p, fvp_gmres_stats *st, bool x_zero) {
    const int m = p->dim_krylov;
    if (m < 1 || m > 63) { set_error("dim_krylov must be 1..63"); return -1; }
    if (p->reorthogonalise < 0 || p->reorthogonalise > 2) { set_error("reorthogonalise must be 0, 1 or 2"); return -1; }
    if (ensure_work(W, m)) return -1;
    GmresWork wk{W->Q.p, W->w.p, W->r.p, W->z.p, W->part.p, W->scal.p, W->hdev.p, W->Hd.p, W->gd.p, W->csd.p,
                 W->snd.p, W->yd.p, W->hstat, W->ev.data(), true, W->hstat + 2 * (m + 1), W->ev[m]};
    return gmres_core(wk, s, n, A, M, b, x, p, st, x_zero);
}

static int gmres_core(const GmresWork &W, hipStream_t s, long n, const Op &A, const Op &M, const double *b,
                      double *x, const fvp_gmres_params *p, fvp_gmres_stats *st, bool x_zero) {
    const int m = p->dim_krylov;
    double *Q = W.Q, *w = W.w, *r = W.r, *z = W.z, *part = W.part, *sc = W.sc;
    // h1, h2: first / second pass coefficients (j+1 dots, then ||w||^2); stat: {estimate, h_{j+1,j}}
    double *h1 = W.hdev, *h2 = W.hdev + (m + 2), *stat = W.hdev + 2 * (m + 2);
    // device scalars: sc[0] ||r||^2, sc[1] ||b||^2, sc[2] first-pass ||w||^2 after projection,
    // sc[3] second-pass ||w||^2, sc[4] reorthogonalisation flag
    double *n1 = sc + 2, *n2 = sc + 3, *flag = sc + 4;
    const long ldq = gs_ldq(n);
    const int mode = p->reorthogonalise;
    fvp_gmres_stats stats{0, 0, 0, 0., 0.};
    double hn2[2] = {0., 0.};
    // r = b - A x; ||r||, ||b|| with one synchronisation
    if (x_zero) {
        vcopy<<<nblk(n, 256), 256, 0, s>>>(n, b, r);
    } else {
        if (A(x, r)) return -1;
        vrsub<<<nblk(n, 256), 256, 0, s>>>(n, b, r);
    }
    // ||r||^2 and ||b||^2 to the host (pinned slots and a polled event when
    // there are, a copy and a blocking wait otherwise)
    auto fetch = [&](int cnt) -> int {
        if (W.hscal) {
            if (host_wait(s, W.hev)) return -1;
            for (int k = 0; k < cnt; ++k) hn2[k] = W.hscal[k];
            return 0;
        }
        FVP_HIP(hipMemcpyAsync(hn2, sc, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
        FVP_HIP(hipStreamSynchronize(s));
        return 0;
    };
    dot_async(s, n, r, r, part, sc, W.hscal);
    dot_async(s, n, b, b, part, sc + 1, W.hscal ? W.hscal + 1 : nullptr);
    if (fetch(2)) return -1;
    double beta = std::sqrt(hn2[0]), normb = std::sqrt(hn2[1]);
    double error = normb > 0 ? beta / normb : 0.;
    if (beta == 0. || error < p->tol) {
        stats.converged = 1; stats.residual = beta; stats.error = error;
        if (st) *st = stats;
        return 0;
    }
    int restarts = 0, converged = 0, total = 0;
    vscale_inv_norm<<<nblk(n, 256), 256, 0, s>>>(n, r, sc, Q);  // q0 = r / beta
    while (restarts <= p->max_restarts && !converged) {
        gmres_cycle_init<<<1, 64, 0, s>>>(W.gd, m + 1, beta);
        int k = m;
        for (int j = 0; j <= m; ++j) {
            if (j < m) {  // queue Arnoldi step j
                const double *qj = Q + (size_t)j * ldq;
                if (M) { if (M(qj, z)) return -1; if (A(z, w)) return -1; }
                else if (A(qj, w)) return -1;
                // first pass: h1 = Q^T w, ||w||^2 -> h1[j+1]; w -= Q h1, ||w||^2 -> n1
                multidot(s, n, w, Q, ldq, j + 1, part, h1, nullptr);
                if (mode && j + 1 <= kChunk) {
                    // first pass's update fused with the second pass's dots (h2, and
                    // ||w||^2 after the first pass into n1 = h2[j + 1]); the second
                    // update only when the first pass cancelled (flag)
                    kAxpyMdot[j](s, gs_blocks(n), n, w, Q, ldq, h1, part, nullptr);
                    reduce_cols<<<j + 2, kThreads, 0, s>>>(gs_blocks(n), part, h2, 0, nullptr,
                                                           mode == 2 ? nullptr : h1 + j + 1, flag);
                    multiaxpy(s, n, w, Q, ldq, j + 1, h2, part, n2, flag);
                } else {
                    // single pass: arnoldi_finish reduces ||w||^2 itself
                    multiaxpy(s, n, w, Q, ldq, j + 1, h1, part, mode ? n1 : nullptr, nullptr);
                    if (mode) {
                        reorth_decide<<<1, 64, 0, s>>>(h1 + j + 1, n1, mode, flag);
                        multidot(s, n, w, Q, ldq, j + 1, part, h2, flag);
                        multiaxpy(s, n, w, Q, ldq, j + 1, h2, part, n2, flag);
                    }
                }
                const bool fused = mode && j + 1 <= kChunk;  // ||w||^2 after the first pass in h2[j + 1]
                arnoldi_finish<<<std::min(nblk(n, 256), kFinishBlocks), 256, 0, s>>>(
                    j, m, h1, h2, fused ? h2 + j + 1 : n1, n2, mode ? flag : nullptr, W.Hd, W.csd, W.snd, W.gd, stat,
                    W.hmapped ? W.hstat + 2 * j : nullptr, n, w, Q + (size_t)(j + 1) * ldq, mode ? nullptr : part,
                    gs_blocks(n));
                if (!W.hmapped)
                    FVP_HIP(hipMemcpyAsync(W.hstat + 2 * j, stat, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
                if (W.ev) FVP_HIP(hipEventRecord(W.ev[j], s));
                else FVP_HIP(hipStreamSynchronize(s));
                FVP_HIP(hipGetLastError());
            }
            if (j > 0) {  // read step j-1 (while step j runs, when pipelined)
                if (W.ev) {  // polled: a blocking wait would wake up late
                    hipError_t e;
                    while ((e = hipEventQuery(W.ev[j - 1])) == hipErrorNotReady) {
                    }
                    FVP_HIP(e);
                }
                ++total;
                const double est = W.hstat[2 * (j - 1)], hn = W.hstat[2 * (j - 1) + 1];
                if (est < p->tol * normb || !(hn > 1e-16)) { k = j; break; }
            }
        }
        gmres_backsolve<<<1, 64, 0, s>>>(k, m, W.Hd, W.gd, W.yd);
        lincomb<<<nblk(n, 256), 256, 0, s>>>(n, w, Q, ldq, k, W.yd);
        if (M) { if (M(w, z)) return -1; }
        else vcopy<<<nblk(n, 256), 256, 0, s>>>(n, w, z);
        vadd<<<nblk(n, 256), 256, 0, s>>>(n, z, x);
        if (A(x, r)) return -1;
        vrsub<<<nblk(n, 256), 256, 0, s>>>(n, b, r);
        dot_async(s, n, r, r, part, sc, W.hscal);
        if (fetch(1)) return -1;
        beta = std::sqrt(hn2[0]);
        error = beta / normb;
        if (error < p->tol) converged = 1;
        else ++restarts;
        if (!converged && restarts <= p->max_restarts)
            vscale_inv_norm<<<nblk(n, 256), 256, 0, s>>>(n, r, sc, Q);
    }
    stats.restarts = converged ? restarts : restarts - 1;
    stats.iterations = total;
    stats.converged = converged;
    stats.residual = beta;
    stats.error = error;
    if (st) *st = stats;
    return converged ? 0 : FVP_ERR_NOCONV;
}


// Restarted GMRES with deflated restarts (gmresFH, gmres.c:212-405, with
// arnoldi, :412-511, and deflate, :527-849): each cycle runs the Arnoldi
// process for A M from column K to m (no early exit within a cycle, as there),
// solves the least-squares problem min || Q^T r - Hbar y || on the host, sets
// x += M Q y and recomputes the true residual; before the next cycle the k
// harmonic Ritz vectors of smallest magnitude (real and imaginary parts of a
// complex pair) of H = H_m + h_{m+1,m}^2 f e_m^T, f = H_m^-T e_m, are kept at
// the front of the basis, with the matching Hbar block, and the next cycle
// continues from column K (the number kept).  The Gram-Schmidt passes, the
// products and the basis updates run on the device; the small dense problems
// on the host.  k = 0 is plain restarted GMRES with convergence checked per
// cycle.
int gmres_deflated(fvp_csr *Wc, hipStream_t s, long n, const Op &A, const Op &M, const double *b, double *x,
                   const fvp_gmres_params *p, int k, fvp_gmres_stats *st) {
    const int m = p->dim_krylov;
    if (m < 2 || m > 63) { set_error("dim_krylov must be 2..63"); return -1; }
    if (k < 0 || 2 * k >= m) { set_error("deflation count k must be 0 <= 2k < dim_krylov"); return -1; }
    if (p->reorthogonalise < 0 || p->reorthogonalise > 2) { set_error("reorthogonalise must be 0, 1 or 2"); return -1; }
    if (ensure_work(Wc, m)) return -1;
    const long ldq = gs_ldq(n);
    if (k && Wc->Qk.n < (size_t)ldq * (2 * k + 1) && Wc->Qk.alloc((size_t)ldq * (2 * k + 1))) return -1;
    double *Q = Wc->Q.p, *w = Wc->w.p, *r = Wc->r.p, *z = Wc->z.p, *part = Wc->part.p, *sc = Wc->scal.p;
    double *h1 = Wc->hdev.p, *h2 = Wc->hdev.p + (m + 2), *coef = Wc->hdev.p;
    double *n1 = sc + 2, *n2 = sc + 3;
    const int nb = nblk(n, 256);
    auto host = [&](const double *d, double *hp, int cnt) -> int {
        FVP_HIP(hipMemcpyAsync(hp, d, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
        FVP_HIP(hipStreamSynchronize(s));
        return 0;
    };
    auto norm = [&](const double *v, double &out) -> int {
        if (dot_async(s, n, v, v, part, sc)) return -1;
        double q;
        if (host(sc, &q, 1)) return -1;
        out = std::sqrt(q);
        return 0;
    };
    fvp_gmres_stats stats{0, 0, 0, 0., 0.};
    // r = b - A x
    if (A(x, r)) return -1;
    vrsub<<<nb, 256, 0, s>>>(n, b, r);
    double beta, normb;
    if (norm(r, beta) || norm(b, normb)) return -1;
    double error = normb > 0 ? beta / normb : 0.;
    if (beta == 0. || error < p->tol) {
        stats.converged = 1; stats.residual = beta; stats.error = error;
        if (st) *st = stats;
        return 0;
    }
    vaxpby<<<nb, 256, 0, s>>>(n, 1. / beta, r, 0., Q);
    std::vector<double> H((size_t)(m + 1) * m, 0.);  // row-major (m+1) x m
    auto Hr = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
    std::vector<double> hh(2 * (m + 2));
    int K = 0, restarts = 0, total = 0, converged = 0;
    while (restarts <= p->max_restarts) {
        // Arnoldi from column K
        int j = K;
        while (j < m) {
            const double *qj = Q + (size_t)j * ldq;
            if (M) { if (M(qj, z)) return -1; if (A(z, w)) return -1; }
            else if (A(qj, w)) return -1;
            multidot(s, n, w, Q, ldq, j + 1, part, h1, nullptr);
            multiaxpy(s, n, w, Q, ldq, j + 1, h1, part, n1, nullptr);
            if (p->reorthogonalise) {
                multidot(s, n, w, Q, ldq, j + 1, part, h2, nullptr);
                multiaxpy(s, n, w, Q, ldq, j + 1, h2, part, n2, nullptr);
            }
            if (host(Wc->hdev.p, hh.data(), 2 * (m + 2)) || host(sc + 2, hh.data() + 2 * (m + 2) - 2, 2)) return -1;
            for (int i = 0; i <= j; ++i) Hr(i, j) = hh[i] + (p->reorthogonalise ? hh[m + 2 + i] : 0.);
            const double ww = hh[2 * (m + 2) - 2 + (p->reorthogonalise ? 1 : 0)];
            const double hn = std::sqrt(std::max(ww, 0.));
            Hr(j + 1, j) = hn;
            ++total;
            if (hn < 1e-16) break;  // lucky breakdown: column j + 1 is not formed (gmres.c:485)
            vaxpby<<<nb, 256, 0, s>>>(n, 1. / hn, w, 0., Q + (size_t)(j + 1) * ldq);
            ++j;
        }
        // least squares on the j computed columns, right-hand side Q^T r
        multidot(s, n, r, Q, ldq, j + 1, part, h1, nullptr);
        std::vector<double> c(j + 1), y, Hs((size_t)(j + 1) * j);
        if (host(h1, c.data(), j + 1)) return -1;
        for (int a = 0; a <= j; ++a)
            for (int bcol = 0; bcol < j; ++bcol) Hs[(size_t)a * j + bcol] = Hr(a, bcol);
        if (j == 0 || !dense::lstsq(j + 1, j, Hs, c, y)) { set_error("GMRES: singular Hessenberg matrix"); return -1; }
        FVP_HIP(hipMemcpyAsync(coef, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, s));
        lincomb<<<nb, 256, 0, s>>>(n, w, Q, ldq, j, coef);
        if (M) { if (M(w, z)) return -1; }
        else vcopy<<<nb, 256, 0, s>>>(n, w, z);
        vadd<<<nb, 256, 0, s>>>(n, z, x);
        if (A(x, r)) return -1;
        vrsub<<<nb, 256, 0, s>>>(n, b, r);
        if (norm(r, beta)) return -1;
        error = beta / normb;
        if (error < p->tol) { converged = 1; break; }
        if (++restarts > p->max_restarts) break;
        int Knew = 0;
        if (k && j == m) {
            // harmonic Ritz vectors of H = H_m + hb^2 f e_m^T (deflate, gmres.c:600-760)
            const double hb = Hr(m, m - 1);
            std::vector<double> Hm((size_t)m * m), HmT((size_t)m * m), f(m, 0.);
            for (int a = 0; a < m; ++a)
                for (int bcol = 0; bcol < m; ++bcol) {
                    Hm[(size_t)a * m + bcol] = Hr(a, bcol);
                    HmT[(size_t)bcol * m + a] = Hr(a, bcol);
                }
            f[m - 1] = 1.;
            bool ok = dense::lu_solve(m, HmT, f);
            std::vector<double> Hh = Hm, wr, wi;
            if (ok)
                for (int a = 0; a < m; ++a) Hh[(size_t)a * m + m - 1] += hb * hb * f[a];
            ok = ok && dense::eig(m, Hh, wr, wi);
            std::vector<int> ord(m);
            std::iota(ord.begin(), ord.end(), 0);
            if (ok)
                std::stable_sort(ord.begin(), ord.end(), [&](int a, int bcol) {
                    return wr[a] * wr[a] + wi[a] * wi[a] < wr[bcol] * wr[bcol] + wi[bcol] * wi[bcol];
                });
            std::vector<double> X;  // column-major m x jj
            for (int i = 0, q = 0; ok && i < k; ++i) {
                const int pos = ord[i + q];
                std::vector<std::complex<double>> v;
                if (wi[pos] != 0.) {
                    // the member of the pair with positive imaginary part: real and imaginary parts
                    const std::complex<double> lam(wr[pos], std::fabs(wi[pos]));
                    ok = dense::eigvec(m, Hh, lam, v);
                    for (int a = 0; a < m; ++a) X.push_back(v[a].real());
                    for (int a = 0; a < m; ++a) X.push_back(v[a].imag());
                    ++q;
                } else {
                    ok = dense::eigvec(m, Hh, std::complex<double>(wr[pos], 0.), v);
                    for (int a = 0; a < m; ++a) X.push_back(v[a].real());
                }
            }
            const int jj = (int)(X.size() / m);
            ok = ok && jj > 0 && jj + 1 <= m && dense::orthonormalise(m, jj, X);
            if (ok) {
                // Qk = Q_m X on the device; p = hb q_{m+1} - hb^2 Q_m f, orthogonalised against Qk
                double *Qk = Wc->Qk.p;
                for (int cc = 0; cc < jj; ++cc) {
                    FVP_HIP(hipMemcpyAsync(coef, &X[(size_t)cc * m], sizeof(double) * m, hipMemcpyHostToDevice, s));
                    lincomb<<<nb, 256, 0, s>>>(n, Qk + (size_t)cc * ldq, Q, ldq, m, coef);
                    FVP_HIP(hipStreamSynchronize(s));  // coef is reused by the next column
                }
                std::vector<double> pc(m + 1);
                for (int a = 0; a < m; ++a) pc[a] = -hb * hb * f[a];
                pc[m] = hb;
                FVP_HIP(hipMemcpyAsync(coef, pc.data(), sizeof(double) * (m + 1), hipMemcpyHostToDevice, s));
                double *pm = Qk + (size_t)jj * ldq;
                lincomb<<<nb, 256, 0, s>>>(n, pm, Q, ldq, m + 1, coef);
                for (int cc = 0; cc < jj; ++cc) {  // modified Gram-Schmidt
                    double dot;
                    if (dot_async(s, n, pm, Qk + (size_t)cc * ldq, part, sc) || host(sc, &dot, 1)) return -1;
                    vaxpby<<<nb, 256, 0, s>>>(n, -dot, Qk + (size_t)cc * ldq, 1., pm);
                }
                double betak;
                if (norm(pm, betak)) return -1;
                ok = betak > 0.;
                if (ok) {
                    vaxpby<<<nb, 256, 0, s>>>(n, 1. / betak, pm, 0., pm);
                    FVP_HIP(hipMemcpyAsync(Q, Qk, sizeof(double) * ldq * (jj + 1), hipMemcpyDeviceToDevice, s));
                    // the new Hbar: Omega = X^T H X - hb^2 (X^T f) s^T, row jj = betak s^T, s = X's last row
                    std::vector<double> HX((size_t)m * jj, 0.), Xf(jj, 0.);
                    for (int a = 0; a < m; ++a)
                        for (int cc = 0; cc < jj; ++cc) {
                            double acc = 0.;
                            for (int q2 = 0; q2 < m; ++q2) acc += Hh[(size_t)a * m + q2] * X[(size_t)cc * m + q2];
                            HX[(size_t)a * jj + cc] = acc;
                        }
                    for (int cc = 0; cc < jj; ++cc)
                        for (int a = 0; a < m; ++a) Xf[cc] += X[(size_t)cc * m + a] * f[a];
                    std::fill(H.begin(), H.end(), 0.);
                    for (int r1 = 0; r1 < jj; ++r1)
                        for (int cc = 0; cc < jj; ++cc) {
                            double om = 0.;
                            for (int a = 0; a < m; ++a) om += X[(size_t)r1 * m + a] * HX[(size_t)a * jj + cc];
                            Hr(r1, cc) = om - hb * hb * Xf[r1] * X[(size_t)cc * m + m - 1];
                        }
                    for (int cc = 0; cc < jj; ++cc) Hr(jj, cc) = betak * X[(size_t)cc * m + m - 1];
                    Knew = jj;
                }
            }
        }
        if (!Knew) {  // plain restart from the residual
            std::fill(H.begin(), H.end(), 0.);
            vaxpby<<<nb, 256, 0, s>>>(n, 1. / beta, r, 0., Q);
        }
        K = Knew;
        FVP_HIP(hipGetLastError());
    }
    stats.restarts = converged ? restarts : restarts - 1;
    stats.iterations = total;
    stats.converged = converged;
    stats.residual = beta;
    stats.error = error;
    if (st) *st = stats;
    return converged ? 0 : FVP_ERR_NOCONV;
}

}  // namespace fvp

// ================================================================================
// decomp-style CSR entry points
// ================================================================================
using namespace fvp;

namespace fvp {

static int csr_gemv_impl(fvp_csr *A, double alpha, const double *x, double beta, double *y) {
    const int L = 8;
    csr_gemv_kernel<L><<<nblk((long)A->n * L, 256), 256, 0, A->stream>>>(A->n, A->ia.p, A->ja.p, A->a.p, alpha, x,
                                                                         beta, y);
    FVP_HIP(hipGetLastError());
    return 0;
}

// level sets of the L (forward) and U (backward) dependency DAGs
static int build_levels(fvp_csr *A, const std::vector<int> &diag) {
    const int n = A->n;
    std::vector<int> lev(n, 0), ulev(n, 0);
    int nl = 0, nu = 0;
    for (int i = 0; i < n; ++i) {
        int l = 0;
        for (int k = A->ia_h[i]; k < diag[i]; ++k) l = std::max(l, lev[A->ja_h[k]] + 1);
        lev[i] = l;
        nl = std::max(nl, l + 1);
    }
    for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        for (int k = diag[i] + 1; k < A->ia_h[i + 1]; ++k) l = std::max(l, ulev[A->ja_h[k]] + 1);
        ulev[i] = l;
        nu = std::max(nu, l + 1);
    }
    auto group = [&](const std::vector<int> &lv, int nlev, std::vector<int> &ptr, std::vector<int> &rows) {
        ptr.assign(nlev + 1, 0);
        for (int i = 0; i < n; ++i) ptr[lv[i] + 1]++;
        for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
        rows.resize(n);
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int i = 0; i < n; ++i) rows[fill[lv[i]]++] = i;
    };
    std::vector<int> lrows, urows;
    group(lev, nl, A->lptr_h, lrows);
    group(ulev, nu, A->uptr_h, urows);
    if (A->lrows.upload(lrows.data(), n) || A->urows.upload(urows.data(), n)) return -1;
    A->levels = true;
    return 0;
}

// Block-local CSR and the per-block level sets of L and U.  Block b owns rows
// [b bs, (b+1) bs); with overlap > 0 its local rows are the owned rows plus
// `overlap` layers of their graph neighbours (sorted by global row, so the
// local ILU(0) follows the global row order), and only the owned rows of its
// solve are kept (restricted additive Schwarz).
static int build_bj(fvp_csr *A) {
    const int n = A->n, BS = A->bs;
    const bool given = !A->bstart.empty();
    const int nb = given ? (int)A->bstart.size() - 1 : (n + BS - 1) / BS;
    std::vector<int> roff(nb + 1, 0), rowid, own(2 * (size_t)nb), bia(1, 0), bsrc, bdiag;
    std::vector<unsigned short> bja;
    std::vector<int> stamp(n, -1), loc(n, 0), ext, frontier, next;
    for (int b = 0; b < nb; ++b) {
        const int lo = given ? A->bstart[b] : b * BS, hi = given ? A->bstart[b + 1] : std::min(n, lo + BS);
        ext.clear();
        frontier.clear();
        for (int i = lo; i < hi; ++i) { stamp[i] = b; ext.push_back(i); frontier.push_back(i); }
        for (int layer = 0; layer < A->overlap; ++layer) {
            next.clear();
            for (int i : frontier)
                for (int k = A->ia_h[i]; k < A->ia_h[i + 1]; ++k) {
                    const int j = A->ja_h[k];
                    if (stamp[j] != b) { stamp[j] = b; ext.push_back(j); next.push_back(j); }
                }
            frontier.swap(next);
        }
        std::sort(ext.begin(), ext.end());
        const int cnt = (int)ext.size();
        if (cnt > 65535) { set_error("block-Jacobi ILU(0): a block has more than 65535 rows"); return -1; }
        for (int r = 0; r < cnt; ++r) loc[ext[r]] = r;
        const int olo = (int)(std::lower_bound(ext.begin(), ext.end(), lo) - ext.begin());
        own[2 * (size_t)b] = olo;
        own[2 * (size_t)b + 1] = olo + (hi - lo);
        for (int r = 0; r < cnt; ++r) {
            const int i = ext[r];
            int d = -1;
            for (int k = A->ia_h[i]; k < A->ia_h[i + 1]; ++k) {
                const int j = A->ja_h[k];
                if (stamp[j] != b) continue;
                if (j == i) d = (int)bsrc.size();
                bsrc.push_back(k);
                bja.push_back((unsigned short)loc[j]);
            }
            if (d < 0) { set_error("zero diagonal on row " + std::to_string(i)); return FVP_ERR_PIVOT; }
            bdiag.push_back(d);
            bia.push_back((int)bsrc.size());
            rowid.push_back(i);
        }
        roff[b + 1] = roff[b] + cnt;
    }
    const int total = roff[nb];
    std::vector<int> blev_l(nb + 1, 0), blev_u(nb + 1, 0), lptr_l(1, 0), lptr_u(1, 0);
    std::vector<unsigned short> rows_l, rows_u;
    std::vector<int> lev;
    int maxnz = 0, maxrow = 0, maxrows = 0;
    for (int b = 0; b < nb; ++b) {
        const int base = roff[b], cnt = roff[b + 1] - base;
        maxrows = std::max(maxrows, cnt);
        maxnz = std::max(maxnz, bia[base + cnt] - bia[base]);
        for (int r = 0; r < cnt; ++r) maxrow = std::max(maxrow, bia[base + r + 1] - bia[base + r]);
        lev.assign(cnt, 0);
        auto group = [&](bool lower, std::vector<int> &lptr, std::vector<unsigned short> &rows) {
            int nl = 0;
            if (lower) {
                for (int r = 0; r < cnt; ++r) {
                    int i = base + r, l = 0;
                    for (int k = bia[i]; k < bdiag[i]; ++k) l = std::max(l, lev[bja[k]] + 1);
                    lev[r] = l;
                    nl = std::max(nl, l + 1);
                }
            } else {
                for (int r = cnt - 1; r >= 0; --r) {
                    int i = base + r, l = 0;
                    for (int k = bdiag[i] + 1; k < bia[i + 1]; ++k) l = std::max(l, lev[bja[k]] + 1);
                    lev[r] = l;
                    nl = std::max(nl, l + 1);
                }
            }
            std::vector<int> c(nl + 1, 0);
            for (int r = 0; r < cnt; ++r) c[lev[r] + 1]++;
            for (int l = 0; l < nl; ++l) c[l + 1] += c[l];
            const int off = (int)rows.size();
            rows.resize(off + cnt);
            std::vector<int> fill(c.begin(), c.end() - 1);
            for (int r = 0; r < cnt; ++r) rows[off + fill[lev[r]]++] = (unsigned short)r;
            for (int l = 0; l < nl; ++l) lptr.push_back(off + c[l + 1]);
            return nl;
        };
        blev_l[b + 1] = blev_l[b] + group(true, lptr_l, rows_l);
        blev_u[b + 1] = blev_u[b] + group(false, lptr_u, rows_u);
    }
    if (maxrow > 64) { set_error("block-Jacobi ILU(0) supports rows of at most 64 block entries"); return -1; }
    if (maxnz > 65535) { set_error("block-Jacobi ILU(0): more than 65535 entries in one block"); return -1; }
    // the factorisation's update pairs (see bj_factor)
    std::vector<unsigned> frec(2 * bsrc.size(), 0u), fpair;
    std::vector<long> fpbase(nb + 1, 0);
    {
        std::vector<int> pos(maxrows, -1);
        for (int b = 0; b < nb; ++b) {
            const int base = roff[b], cnt = roff[b + 1] - base, z0 = bia[base];
            fpbase[b] = (long)fpair.size();
            for (int r = 0; r < cnt; ++r) {
                const int s0 = bia[base + r], s1 = bia[base + r + 1], dd = bdiag[base + r];
                for (int q = s0; q < s1; ++q) pos[bja[q]] = q - z0;
                for (int e = s0; e < dd; ++e) {
                    const int k = bja[e], kd = bdiag[base + k], ke = bia[base + k + 1];
                    const long start = (long)fpair.size() - fpbase[b];
                    unsigned count = 0;
                    for (int m = kd + 1; m < ke; ++m) {
                        const int p = pos[bja[m]];
                        if (p >= 0) { fpair.push_back((unsigned)(m - z0) | ((unsigned)p << 16)); ++count; }
                    }
                    frec[2 * (size_t)e] = (unsigned)start;
                    frec[2 * (size_t)e + 1] = count | ((unsigned)(kd - z0) << 16);
                }
                for (int q = s0; q < s1; ++q) pos[bja[q]] = -1;
            }
        }
        fpbase[nb] = (long)fpair.size();
        if (fpair.empty()) fpair.push_back(0u);  // a block-diagonal matrix has no pairs
    }
    if (A->fac_rec.upload(frec.data(), frec.size()) || A->fac_pair.upload(fpair.data(), fpair.size()) ||
        A->fac_pbase.upload(fpbase.data(), nb + 1))
        return -1;
    const bool ident = A->overlap == 0;  // rows of block b are [b bs, ...): no row list needed
    if (A->b_ia.upload(bia.data(), total + 1) || A->b_src.upload(bsrc.data(), bsrc.size()) ||
        A->b_diag.upload(bdiag.data(), total) || A->b_ja.upload(bja.data(), bja.size()) ||
        A->b_lu.alloc(bsrc.size()) || A->b_roff.upload(roff.data(), nb + 1) ||
        A->b_own.upload(own.data(), own.size()) || (!ident && A->b_rowid.upload(rowid.data(), total)) ||
        A->b_lev_l.upload(blev_l.data(), nb + 1) || A->b_lev_u.upload(blev_u.data(), nb + 1) ||
        A->b_lptr_l.upload(lptr_l.data(), lptr_l.size()) || A->b_lptr_u.upload(lptr_u.data(), lptr_u.size()) ||
        A->b_rows_l.upload(rows_l.data(), rows_l.size()) || A->b_rows_u.upload(rows_u.data(), rows_u.size()))
        return -1;
    if (ident) A->b_rowid.release();
    A->b_lu32.release();
    {
        // packed solve layout (see bj_solve)
        std::vector<int> ovp(nb + 1, 0);
        std::vector<long> ovpos;
        std::vector<uint8_t> ovtri;
        std::vector<int> dsr;
        struct Tri {
            std::vector<unsigned> meta;
            std::vector<long> ebase;
            std::vector<int> esrc;
            std::vector<unsigned short> col;
        } tr[2];
        for (int t = 0; t < 2; ++t) tr[t].ebase.assign(nb + 1, 0);
        int ovmax = 0;
        for (int b = 0; b < nb; ++b) {
            const int base = roff[b];
            int nov = 0;
            for (int tri = 0; tri < 2; ++tri) {
                Tri &T = tr[tri];
                const long eb = (long)T.esrc.size();
                T.ebase[b] = eb;
                const std::vector<int> &lptr = tri ? lptr_u : lptr_l;
                const std::vector<int> &blev = tri ? blev_u : blev_l;
                const std::vector<unsigned short> &rows = tri ? rows_u : rows_l;
                for (int sl = lptr[blev[b]]; sl < lptr[blev[b + 1]]; ++sl) {
                    const int r = rows[sl];
                    const int i = base + r;
                    const int k0 = tri ? bdiag[i] + 1 : bia[i], k1 = tri ? bia[i + 1] : bdiag[i];
                    const int kn = k1 - k0;
                    const long kl = (long)T.esrc.size() - eb;
                    if (kl > 65535 || nov > 65535) {
                        set_error("block-Jacobi ILU(0): more than 65535 entries in one block triangle");
                        return -1;
                    }
                    T.meta.push_back((unsigned)kl | ((unsigned)nov << 16));
                    T.meta.push_back((unsigned)r | ((unsigned)kn << 16));
                    for (int k = k0; k < k1; ++k) {
                        if (k - k0 >= kBjRow) {
                            ovpos.push_back((long)T.esrc.size());
                            ovtri.push_back((uint8_t)tri);
                            ++nov;
                        }
                        T.esrc.push_back(k);
                        T.col.push_back(bja[k]);
                    }
                    if (kn & 1) { T.esrc.push_back(-1); T.col.push_back((unsigned short)r); }
                    if (tri) dsr.push_back(bdiag[i]);
                }
            }
            ovp[b + 1] = ovp[b] + nov;
            ovmax = std::max(ovmax, nov);
        }
        for (int t = 0; t < 2; ++t) {  // the last windows read up to kBjRow entries past the end
            tr[t].ebase[nb] = (long)tr[t].esrc.size();
            for (int k = 0; k < kBjRow; ++k) { tr[t].esrc.push_back(-1); tr[t].col.push_back(0); }
        }
        // the solve's levels: each factor level cut into pieces of at most 32 slots (one per lane pair)
        auto split = [&](const std::vector<int> &blev, const std::vector<int> &lptr, std::vector<int> &sb,
                         std::vector<int> &sp, int &levmax) {
            sb.assign(nb + 1, 0);
            sp.assign(1, 0);
            levmax = 0;
            for (int b = 0; b < nb; ++b) {
                for (int l = blev[b]; l < blev[b + 1]; ++l)
                    for (int p = lptr[l]; p < lptr[l + 1]; p += 32) sp.push_back(std::min(p + 32, lptr[l + 1]));
                sb[b + 1] = (int)sp.size() - 1;
                levmax = std::max(levmax, sb[b + 1] - sb[b]);
            }
        };
        std::vector<int> sbl, spl, sbu, spu;
        split(blev_l, lptr_l, sbl, spl, A->s_levmax_l);
        split(blev_u, lptr_u, sbu, spu, A->s_levmax_u);
        A->s_maxslots = 0;
        for (int b = 0; b < nb; ++b)
            A->s_maxslots = std::max({A->s_maxslots, spl[sbl[b + 1]] - spl[sbl[b]], spu[sbu[b + 1]] - spu[sbu[b]]});
        if (A->meta_l.upload(tr[0].meta.data(), tr[0].meta.size()) ||
            A->meta_u.upload(tr[1].meta.data(), tr[1].meta.size()) ||
            A->ebase_l.upload(tr[0].ebase.data(), nb + 1) || A->ebase_u.upload(tr[1].ebase.data(), nb + 1) ||
            A->esrc_l.upload(tr[0].esrc.data(), tr[0].esrc.size()) ||
            A->esrc_u.upload(tr[1].esrc.data(), tr[1].esrc.size()) ||
            A->col_l.upload(tr[0].col.data(), tr[0].col.size()) || A->col_u.upload(tr[1].col.data(), tr[1].col.size()) ||
            A->dsrc.upload(dsr.data(), dsr.size()) || A->ov_ptr.upload(ovp.data(), nb + 1) ||
            A->ov_pos.upload(ovpos.data(), ovpos.size()) || A->ov_tri.upload(ovtri.data(), ovtri.size()) ||
            A->s_lev_l.upload(sbl.data(), nb + 1) || A->s_lev_u.upload(sbu.data(), nb + 1) ||
            A->s_lptr_l.upload(spl.data(), spl.size()) || A->s_lptr_u.upload(spu.data(), spu.size()))
            return -1;
        A->ov_max = ovmax;
        A->val_l.release(); A->val_u.release(); A->dinv.release();
        A->val32_l.release(); A->val32_u.release(); A->dinv32.release();
    }
    A->b_maxnz = maxnz;
    A->b_maxrows = maxrows;
    if (hipDeviceGetAttribute(&A->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, A->device) != hipSuccess)
        A->lds_max = 64 * 1024;
    // per-kernel attributes (shared by every matrix): allow the device maximum;
    // each build checks its own block against it
    for (const void *f : {(const void *)bj_factor<double, true>, (const void *)bj_factor<double, false>,
                          (const void *)bj_solve<double, false>, (const void *)bj_solve<double, true>,
                          (const void *)bj_factor<float, true>, (const void *)bj_factor<float, false>,
                          (const void *)bj_solve<float, false>, (const void *)bj_solve<float, true>})
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, A->lds_max);
    A->nblk = nb;
    A->bj = true;
    return 0;
}

int csr_create_impl(int device, int n, const int *ia, const int *ja, fvp_csr **out);
static int precon_build_impl(fvp_csr *A, int type);
static int precon_apply_impl(fvp_csr *A, const double *x, double *y, const double *xdiv = nullptr,
                             double *tmp = nullptr, double *yn2 = nullptr);

static void schur_free(fvp_csr *A) {
    fvp_schur *P = A->sc;
    if (!P) return;
    if (P->B) fvp_csr_destroy(P->B);
    P->irow.release(); P->grow.release(); P->bsrc.release(); P->e_ia.release(); P->e_ja.release();
    P->e_loc.release(); P->e_src.release(); P->f_ia.release(); P->f_ja.release(); P->f_src.release();
    P->ev.release(); P->fv.release(); P->c_src.release(); P->c_pos.release(); P->goff.release();
    P->gsub.release(); P->glocal.release(); P->soff.release(); P->S.release(); P->t1.release();
    P->t2.release(); P->yG.release();
    delete P;
    A->sc = nullptr;
}

// Split of mtx_CRS_dist_split (schur.c:82-600) for subdomains of bs consecutive
// rows: a row is a boundary row when it has an entry in another subdomain or
// another subdomain has an entry in its column.  Interior unknowns keep their
// row order inside each subdomain, and so do the boundary unknowns.
static int build_schur(fvp_csr *A) {
    const int n = A->n, bs = A->bs, nsub = (n + bs - 1) / bs;
    std::vector<char> bnd(n, 0);
    for (int i = 0; i < n; ++i)
        for (int k = A->ia_h[i]; k < A->ia_h[i + 1]; ++k) {
            const int j = A->ja_h[k];
            if (i / bs != j / bs) bnd[i] = bnd[j] = 1;
        }
    std::vector<int> irow, grow, idx(n), goff(nsub + 1, 0), bstart(1, 0), gsub, glocal;
    for (int b = 0; b < nsub; ++b) {
        const int lo = b * bs, hi = std::min(n, lo + bs);
        for (int i = lo; i < hi; ++i)
            if (!bnd[i]) { idx[i] = (int)irow.size(); irow.push_back(i); }
        if ((int)irow.size() > bstart.back()) bstart.push_back((int)irow.size());  // empty interiors are skipped
        for (int i = lo; i < hi; ++i)
            if (bnd[i]) {
                idx[i] = (int)grow.size();
                glocal.push_back((int)grow.size() - goff[b]);
                grow.push_back(i);
                gsub.push_back(b);
            }
        goff[b + 1] = (int)grow.size();
        if (goff[b + 1] - goff[b] > kSchurMax) {
            set_error("Schur preconditioner: subdomain " + std::to_string(b) + " has " +
                      std::to_string(goff[b + 1] - goff[b]) + " boundary rows, at most " +
                      std::to_string(kSchurMax) + " are supported (use smaller subdomains)");
            return -1;
        }
    }
    const int nI = (int)irow.size(), nG = (int)grow.size();
    std::vector<int> bia(1, 0), bja, bsrc, eia(1, 0), eja, eloc, esrc, fia(1, 0), fja, fsrc, csrc;
    std::vector<long> cpos, soff(nsub + 1, 0);
    for (int b = 0; b < nsub; ++b) {
        const long m = goff[b + 1] - goff[b];
        soff[b + 1] = soff[b] + m * m;
    }
    for (int i : irow) {
        for (int k = A->ia_h[i]; k < A->ia_h[i + 1]; ++k) {
            const int j = A->ja_h[k];
            if (bnd[j]) { eja.push_back(idx[j]); eloc.push_back(idx[j] - goff[j / bs]); esrc.push_back(k); }
            else { bja.push_back(idx[j]); bsrc.push_back(k); }
        }
        bia.push_back((int)bja.size());
        eia.push_back((int)eja.size());
    }
    for (int i : grow) {
        const int b = i / bs;
        for (int k = A->ia_h[i]; k < A->ia_h[i + 1]; ++k) {
            const int j = A->ja_h[k];
            if (!bnd[j]) { fja.push_back(idx[j]); fsrc.push_back(k); }
            else if (j / bs == b) {  // C_ii; couplings to other subdomains' boundary rows are dropped
                const long m = goff[b + 1] - goff[b];
                cpos.push_back(soff[b] + (long)(idx[j] - goff[b]) * m + (idx[i] - goff[b]));
                csrc.push_back(k);
            }
        }
        fia.push_back((int)fja.size());
    }
    schur_free(A);
    fvp_schur *P = new fvp_schur();
    A->sc = P;
    P->nI = nI; P->nG = nG; P->nsub = nsub;
    P->maxg = 0;
    for (int b = 0; b < nsub; ++b) P->maxg = std::max(P->maxg, goff[b + 1] - goff[b]);
    if (nI > 0) {
        if (csr_create_impl(A->device, nI, bia.data(), bja.data(), &P->B)) { schur_free(A); return -1; }
        P->B->bstart = bstart;
        if (P->irow.upload(irow.data(), nI) || P->bsrc.upload(bsrc.data(), bsrc.size()) ||
            P->e_ia.upload(eia.data(), nI + 1) || P->t1.alloc(nI) || P->t2.alloc(nI) ||
            (!eja.empty() && (P->e_ja.upload(eja.data(), eja.size()) || P->e_loc.upload(eloc.data(), eloc.size()) ||
                              P->e_src.upload(esrc.data(), esrc.size()) || P->ev.alloc(eja.size())))) {
            schur_free(A);
            return -1;
        }
    }
    if (nG > 0) {
        if (P->grow.upload(grow.data(), nG) || P->f_ia.upload(fia.data(), nG + 1) || P->yG.alloc(nG) ||
            P->gsub.upload(gsub.data(), nG) || P->glocal.upload(glocal.data(), nG) ||
            P->goff.upload(goff.data(), nsub + 1) || P->soff.upload(soff.data(), nsub + 1) ||
            P->S.alloc(soff[nsub]) || P->c_src.upload(csrc.data(), csrc.size()) ||
            P->c_pos.upload(cpos.data(), cpos.size()) ||
            (!fja.empty() && (P->f_ja.upload(fja.data(), fja.size()) || P->f_src.upload(fsrc.data(), fsrc.size()) ||
                              P->fv.alloc(fja.size())))) {
            schur_free(A);
            return -1;
        }
    }
    return 0;
}

// precon_schur (schur.c:1247-1323) with schur_form_local (schur.c:887-981):
// ILU(0) of the interior blocks, then S_i = C_ii - F_i M_B^-1 E_i one column
// index at a time for every subdomain at once, then each S_i inverted.
static int schur_factor(fvp_csr *A) {
    hipStream_t s = A->stream;
    if (!A->lds_max && hipDeviceGetAttribute(&A->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, A->device))
        A->lds_max = 64 * 1024;
    if (!A->sc && build_schur(A)) return -1;
    fvp_schur *P = A->sc;
    const size_t lds = (size_t)P->maxg * (P->maxg + 2) * sizeof(double) + (size_t)(P->maxg + 2) * sizeof(int);
    if (lds > (size_t)A->lds_max) { set_error("Schur preconditioner: a dense block does not fit the LDS"); return -1; }
    if (lds > 65536 && hipFuncSetAttribute((const void *)schur_invert, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        set_error("Schur preconditioner: the block inversion cannot have " + std::to_string(lds) + " bytes of LDS");
        return -1;
    }
    if (P->B) {
        fvp_csr *B = P->B;
        fvp_csr_set_stream(B, s);
        vgather<<<nblk(B->nnz, 256), 256, 0, s>>>(B->nnz, P->bsrc.p, A->a.p, B->a.p);
        int rc = precon_build_impl(B, FVP_PRECON_BJILU0);
        if (rc) return rc;
        if (P->ev.n) vgather<<<nblk(P->ev.n, 256), 256, 0, s>>>(P->ev.n, P->e_src.p, A->a.p, P->ev.p);
    }
    if (P->nG == 0) return 0;
    if (P->fv.n) vgather<<<nblk(P->fv.n, 256), 256, 0, s>>>(P->fv.n, P->f_src.p, A->a.p, P->fv.p);
    for (int c = 0; c < P->maxg; ++c) {
        if (P->B) {
            schur_col_rhs<<<nblk(P->nI, 256), 256, 0, s>>>(P->nI, P->e_ia.p, P->e_loc.p, P->ev.p, c, P->t1.p);
            int rc = precon_apply_impl(P->B, P->t1.p, P->t2.p);
            if (rc) return rc;
        }
        schur_col_store<<<nblk(P->nG, 256), 256, 0, s>>>(P->nG, P->f_ia.p, P->f_ja.p, P->fv.p, P->t2.p, P->gsub.p,
                                                        P->glocal.p, P->goff.p, P->soff.p, c, P->S.p);
    }
    if (P->c_src.n)
        schur_add_c<<<nblk(P->c_src.n, 256), 256, 0, s>>>(P->c_src.n, P->c_pos.p, P->c_src.p, A->a.p, P->S.p);
    schur_invert<<<P->nsub, 256, lds, s>>>(P->goff.p, P->soff.p, P->S.p, A->flag.p);
    FVP_HIP(hipGetLastError());
    return 0;
}

// precon_schur_apply (schur.c:1583-1626): u = M_B^-1 x_I; y_G = S^-1 (x_G - F u);
// y_I = M_B^-1 (x_I - E y_G)
static int schur_apply(fvp_csr *A, const double *x, double *y) {
    hipStream_t s = A->stream;
    fvp_schur *P = A->sc;
    if (P->B) {
        fvp_csr_set_stream(P->B, s);
        vgather<<<nblk(P->nI, 256), 256, 0, s>>>(P->nI, P->irow.p, x, P->t1.p);
        int rc = precon_apply_impl(P->B, P->t1.p, P->t2.p);
        if (rc) return rc;
    }
    if (P->nG)
        schur_mid<<<P->nsub, kSchurMax, 0, s>>>(P->goff.p, P->soff.p, P->S.p, P->grow.p, P->f_ia.p, P->f_ja.p,
                                                 P->fv.p, x, P->t2.p, y, P->yG.p);
    if (P->B) {
        schur_back<<<nblk(P->nI, 256), 256, 0, s>>>(P->nI, P->irow.p, P->e_ia.p, P->e_ja.p, P->ev.p, x, P->yG.p,
                                                   P->t1.p);
        int rc = precon_apply_impl(P->B, P->t1.p, P->t2.p);
        if (rc) return rc;
        vscatter<<<nblk(P->nI, 256), 256, 0, s>>>(P->nI, P->irow.p, P->t2.p, y);
    }
    return 0;
}

static int precon_build_impl(fvp_csr *A, int type) {
    hipStream_t s = A->stream;
    const int n = A->n;
    if (type == FVP_PRECON_NONE) { A->ptype = type; return 0; }
    if (type != FVP_PRECON_JACOBI && type != FVP_PRECON_ILU0 && type != FVP_PRECON_BJILU0 &&
        type != FVP_PRECON_BJILU0_F32 && type != FVP_PRECON_SCHUR) {
        set_error("invalid preconditioner type");
        return -1;
    }
    if (!A->flag.p && A->flag.alloc(1)) return -1;
    FVP_HIP(hipMemsetAsync(A->flag.p, 0, sizeof(int), s));
    if (type == FVP_PRECON_SCHUR) {
        int rc = schur_factor(A);
        if (rc) { A->ptype = -1; return rc; }
    } else if (type == FVP_PRECON_BJILU0 || type == FVP_PRECON_BJILU0_F32) {
        if (!A->bj) {
            int rc = build_bj(A);
            if (rc) { A->ptype = -1; return rc; }
        }
        const size_t vb = type == FVP_PRECON_BJILU0 ? sizeof(double) : sizeof(float);
        // values staged in LDS when two workgroups fit a CU (FVP_BJ_FACTOR_MODE 0 / 1
        // forces in place / staged where it fits the workgroup: a tuning and test knob)
        const size_t staged = (size_t)A->b_maxnz * vb;
        bool stage = staged <= (size_t)A->lds_max / 2;
        if (const char *e = getenv("FVP_BJ_FACTOR_MODE")) stage = atoi(e) > 0 && staged <= (size_t)A->lds_max;
        A->factor_mode = stage ? 1 : 0;
        const uint2 *frec = reinterpret_cast<const uint2 *>(A->fac_rec.p);
#define FVP_BJ_FACTOR(TT, ST, LU)                                                                                 \
    bj_factor<TT, ST><<<A->nblk, kBjFactorThreads, ST ? staged : 0, s>>>(                                         \
        A->b_roff.p, A->b_ia.p, A->b_diag.p, A->b_src.p, A->a.p, frec, A->fac_pbase.p, A->fac_pair.p, A->b_lev_l.p, \
        A->b_lptr_l.p, A->b_rows_l.p, LU, A->flag.p)
        if (type == FVP_PRECON_BJILU0) {
            if (stage) FVP_BJ_FACTOR(double, true, A->b_lu.p);
            else FVP_BJ_FACTOR(double, false, A->b_lu.p);
        } else {
            if (!A->b_lu32.p && A->b_lu32.alloc(A->b_lu.n)) return -1;
            if (stage) FVP_BJ_FACTOR(float, true, A->b_lu32.p);
            else FVP_BJ_FACTOR(float, false, A->b_lu32.p);
        }
#undef FVP_BJ_FACTOR
        // packed solve values of the new factor
        const long nl = (long)A->esrc_l.n, nu = (long)A->esrc_u.n, nd = (long)A->dsrc.n;
        if (type == FVP_PRECON_BJILU0) {
            if (!A->val_l.p && (A->val_l.alloc(nl) || A->val_u.alloc(nu) || A->dinv.alloc(nd))) return -1;
            if (nl) bj_pack_vals<double><<<nblk(nl, 256), 256, 0, s>>>(nl, A->esrc_l.p, A->b_lu.p, A->val_l.p);
            if (nu) bj_pack_vals<double><<<nblk(nu, 256), 256, 0, s>>>(nu, A->esrc_u.p, A->b_lu.p, A->val_u.p);
            if (nd) bj_pack_vals<double><<<nblk(nd, 256), 256, 0, s>>>(nd, A->dsrc.p, A->b_lu.p, A->dinv.p);
        } else {
            if (!A->val32_l.p && (A->val32_l.alloc(nl) || A->val32_u.alloc(nu) || A->dinv32.alloc(nd))) return -1;
            if (nl) bj_pack_vals<float><<<nblk(nl, 256), 256, 0, s>>>(nl, A->esrc_l.p, A->b_lu32.p, A->val32_l.p);
            if (nu) bj_pack_vals<float><<<nblk(nu, 256), 256, 0, s>>>(nu, A->esrc_u.p, A->b_lu32.p, A->val32_u.p);
            if (nd) bj_pack_vals<float><<<nblk(nd, 256), 256, 0, s>>>(nd, A->dsrc.p, A->b_lu32.p, A->dinv32.p);
        }
    } else if (type == FVP_PRECON_JACOBI) {
        if (A->jac.n != (size_t)n && A->jac.alloc(n)) return -1;
        jacobi_build<<<nblk(n, 256), 256, 0, s>>>(n, A->ia.p, A->ja.p, A->a.p, A->jac.p, A->flag.p);
    } else {
        if (!A->diag.p) {
            std::vector<int> diag(n);
            for (int i = 0; i < n; ++i) {
                int k = A->ia_h[i];
                while (k < A->ia_h[i + 1] && A->ja_h[k] != i) ++k;
                if (k == A->ia_h[i + 1]) { set_error("zero diagonal on row " + std::to_string(i)); A->ptype = -1; return FVP_ERR_PIVOT; }
                diag[i] = k;
            }
            if (A->diag.upload(diag.data(), n)) return -1;
            if (build_levels(A, diag)) return -1;
        }
        if (A->lu.n != (size_t)A->nnz && A->lu.alloc(A->nnz)) return -1;
        FVP_HIP(hipMemcpyAsync(A->lu.p, A->a.p, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, s));
        for (size_t l = 0; l + 1 < A->lptr_h.size(); ++l) {
            int cnt = A->lptr_h[l + 1] - A->lptr_h[l];
            ilu0_factor_level<<<nblk(cnt, 64), 64, 0, s>>>(cnt, A->lrows.p + A->lptr_h[l], A->ia.p, A->ja.p,
                                                           A->diag.p, A->lu.p, A->flag.p);
        }
    }
    FVP_HIP(hipGetLastError());
    int flag = 0;
    FVP_HIP(hipMemcpyAsync(&flag, A->flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    FVP_HIP(hipStreamSynchronize(s));
    if (flag) { set_error("zero pivot in preconditioner factorisation"); A->ptype = -1; return FVP_ERR_PIVOT; }
    A->ptype = type;
    return 0;
}

// y = M^-1 (x / xdiv) with xdiv optional (the block solves divide as they load);
// yn2 (optional, at least nblk doubles): the block solves' partials of ||y||^2
static int precon_apply_impl(fvp_csr *A, const double *x, double *y, const double *xdiv, double *tmp,
                             double *yn2) {
    hipStream_t s = A->stream;
    const int n = A->n;
    if (xdiv && A->ptype != FVP_PRECON_BJILU0 && A->ptype != FVP_PRECON_BJILU0_F32) {
        if (!tmp) { set_error("internal: scaled preconditioner solve needs a work vector"); return -1; }
        vdiv<<<nblk(n, 256), 256, 0, s>>>(n, x, xdiv, tmp);
        x = tmp;
        xdiv = nullptr;
    }
    switch (A->ptype) {
    case FVP_PRECON_NONE:
        vcopy<<<nblk(n, 256), 256, 0, s>>>(n, x, y);
        break;
    case FVP_PRECON_JACOBI:
        vmul<<<nblk(n, 256), 256, 0, s>>>(n, x, A->jac.p, y);
        break;
    case FVP_PRECON_BJILU0:
    case FVP_PRECON_BJILU0_F32: {
        const bool f64 = A->ptype == FVP_PRECON_BJILU0;
        const size_t ts = f64 ? sizeof(double) : sizeof(float);
        size_t lds = bj_solve_lds(ts, A->b_maxrows, A->ov_max, A->s_levmax_l, A->s_levmax_u, A->s_maxslots);
        const bool ovg = lds > (size_t)A->lds_max;  // overflow lists too long for the LDS: read them in memory
        if (ovg) lds = bj_solve_lds(ts, A->b_maxrows, 0, A->s_levmax_l, A->s_levmax_u, A->s_maxslots);
        if (lds > (size_t)A->lds_max) {
            set_error("block-Jacobi ILU(0): a block's solution does not fit the LDS");
            return -1;
        }
        const uint2 *ml = reinterpret_cast<const uint2 *>(A->meta_l.p), *mu = reinterpret_cast<const uint2 *>(A->meta_u.p);
#define FVP_BJ_SOLVE(TT, OV, VL, VU, DI)                                                                          \
    bj_solve<TT, OV><<<A->nblk, 64, lds, s>>>(A->b_roff.p, A->b_rowid.p, A->b_own.p, A->s_lev_l.p, A->s_lptr_l.p, \
                                              ml, A->ebase_l.p, VL, A->col_l.p, A->s_lev_u.p, A->s_lptr_u.p, mu, \
                                              A->ebase_u.p, VU, A->col_u.p, DI, A->ov_ptr.p, A->ov_pos.p,         \
                                              A->ov_tri.p, A->b_maxrows, A->ov_max, A->s_levmax_l,       \
                                              A->s_levmax_u, x, xdiv, y, yn2)
        if (f64) {
            if (ovg) FVP_BJ_SOLVE(double, true, A->val_l.p, A->val_u.p, A->dinv.p);
            else FVP_BJ_SOLVE(double, false, A->val_l.p, A->val_u.p, A->dinv.p);
        } else {
            if (ovg) FVP_BJ_SOLVE(float, true, A->val32_l.p, A->val32_u.p, A->dinv32.p);
            else FVP_BJ_SOLVE(float, false, A->val32_l.p, A->val32_u.p, A->dinv32.p);
        }
#undef FVP_BJ_SOLVE
        break;
    }
    case FVP_PRECON_SCHUR: {
        int rc = schur_apply(A, x, y);
        if (rc) return rc;
        break;
    }
    case FVP_PRECON_ILU0: {
        const int L = 8;
        for (size_t l = 0; l + 1 < A->lptr_h.size(); ++l) {
            int cnt = A->lptr_h[l + 1] - A->lptr_h[l];
            ilu0_lower_level<L><<<nblk((long)cnt * L, 256), 256, 0, s>>>(cnt, A->lrows.p + A->lptr_h[l], A->ia.p,
                                                                        A->ja.p, A->diag.p, A->lu.p, x, y);
        }
        for (size_t l = 0; l + 1 < A->uptr_h.size(); ++l) {
            int cnt = A->uptr_h[l + 1] - A->uptr_h[l];
            ilu0_upper_level<L><<<nblk((long)cnt * L, 256), 256, 0, s>>>(cnt, A->urows.p + A->uptr_h[l], A->ia.p,
                                                                        A->ja.p, A->diag.p, A->lu.p, y);
        }
        break;
    }
    default:
        set_error("preconditioner not built");
        return -1;
    }
    FVP_HIP(hipGetLastError());
    return 0;
}

int csr_create_impl(int device, int n, const int *ia, const int *ja, fvp_csr **out) {
    if (!out || !ia || !ja || n <= 0) { set_error("invalid CSR arguments"); return -1; }
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (ia[i + 1] < ia[i]) { set_error("row pointers must be non-decreasing"); return -1; }
        for (int k = ia[i]; k < ia[i + 1]; ++k) {
            if (ja[k] < 0 || ja[k] >= n) { set_error("column index out of range"); return -1; }
            if (k > ia[i] && ja[k] <= ja[k - 1]) { set_error("column indices must be sorted and unique"); return -1; }
        }
    }
    FVP_HIP(hipSetDevice(device));
    fvp_csr *A = new fvp_csr();
    A->device = device;
    A->n = n;
    A->nnz = ia[n];
    A->ia_h.assign(ia, ia + n + 1);
    A->ja_h.assign(ja, ja + A->nnz);
    if (hipStreamCreateWithFlags(&A->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete A; set_error("hipStreamCreate failed"); return -1;
    }
    A->stream = A->own_stream;
    if (A->ia.upload(ia, n + 1) || A->ja.upload(ja, A->nnz) || A->a.alloc(A->nnz)) { fvp_csr_destroy(A); return -1; }
    *out = A;
    return 0;
}

}  // namespace fvp

extern "C" int fvp_csr_create(int device, int n, const int *ia, const int *ja, fvp_csr **out) {
    return csr_create_impl(device, n, ia, ja, out);
}

extern "C" int fvp_csr_destroy(fvp_csr *A) {
    if (!A) return 0;
    (void)hipSetDevice(A->device);
    if (A->own_stream) (void)hipStreamSynchronize(A->own_stream);
    A->ia.release(); A->ja.release(); A->a.release(); A->lu.release(); A->jac.release(); A->diag.release();
    A->lrows.release(); A->urows.release(); A->flag.release(); A->Q.release(); A->w.release(); A->r.release();
    A->z.release(); A->part.release(); A->scal.release(); A->hdev.release();
    A->Hd.release(); A->gd.release(); A->csd.release(); A->snd.release(); A->yd.release(); A->Qk.release();
    if (A->hstat) (void)hipHostFree(A->hstat);
    for (hipEvent_t e : A->ev) (void)hipEventDestroy(e);
    A->b_ia.release(); A->b_src.release(); A->b_diag.release(); A->b_ja.release(); A->b_lu.release(); A->b_lu32.release();
    A->b_roff.release(); A->b_rowid.release(); A->b_own.release(); A->b_zn2.release();
    A->meta_l.release(); A->meta_u.release(); A->ebase_l.release(); A->ebase_u.release(); A->esrc_l.release();
    A->esrc_u.release(); A->dsrc.release(); A->col_l.release(); A->col_u.release(); A->val_l.release();
    A->val_u.release(); A->dinv.release(); A->val32_l.release(); A->val32_u.release(); A->dinv32.release();
    A->b_lev_l.release(); A->b_lev_u.release(); A->b_lptr_l.release(); A->b_lptr_u.release();
    A->b_rows_l.release(); A->b_rows_u.release();
    A->fac_rec.release(); A->fac_pair.release(); A->fac_pbase.release();
    A->s_lev_l.release(); A->s_lev_u.release(); A->s_lptr_l.release(); A->s_lptr_u.release();
    A->ov_ptr.release(); A->ov_pos.release(); A->ov_tri.release();
    schur_free(A);
    if (A->own_stream) (void)hipStreamDestroy(A->own_stream);
    delete A;
    return 0;
}

extern "C" int fvp_csr_set_stream(fvp_csr *A, void *stream) {
    if (!A) { set_error("null matrix"); return -1; }
    A->stream = stream ? (hipStream_t)stream : A->own_stream;
    return 0;
}

extern "C" int fvp_csr_set_values(fvp_csr *A, const double *values) {
    if (!A || !values) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    FVP_HIP(hipMemcpyAsync(A->a.p, values, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, A->stream));
    A->ptype = -1;
    return 0;
}

extern "C" int fvp_csr_get_values(fvp_csr *A, double *values_host) {
    if (!A || !values_host) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    FVP_HIP(hipStreamSynchronize(A->stream));
    FVP_HIP(hipMemcpy(values_host, A->a.p, sizeof(double) * A->nnz, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int fvp_csr_gemv(fvp_csr *A, double alpha, const double *x, double beta, double *y) {
    if (!A || !x || !y) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    return csr_gemv_impl(A, alpha, x, beta, y);
}

static int set_block_rows(fvp_csr *A, int rows) {
    if (rows < 64 || rows > 4096 || (rows & (rows - 1))) {
        set_error("block rows must be a power of two in [64, 4096]");
        return -1;
    }
    if (rows != A->bs) {
        A->bs = rows;
        A->bj = false;  // block tables are rebuilt by the next block-Jacobi build
        schur_free(A);  // and the subdomains by the next Schur build
        if (A->ptype == FVP_PRECON_BJILU0 || A->ptype == FVP_PRECON_BJILU0_F32 || A->ptype == FVP_PRECON_SCHUR)
            A->ptype = -1;
    }
    return 0;
}

extern "C" int fvp_precon_block_rows(fvp_csr *A, int rows) {
    if (!A) { set_error("null matrix"); return -1; }
    return set_block_rows(A, rows);
}

static int set_overlap(fvp_csr *A, int layers) {
    if (layers < 0 || layers > 2) { set_error("overlap must be 0, 1 or 2 layers"); return -1; }
    if (layers != A->overlap) {
        A->overlap = layers;
        A->bj = false;
        if (A->ptype == FVP_PRECON_BJILU0 || A->ptype == FVP_PRECON_BJILU0_F32) A->ptype = -1;
    }
    return 0;
}

extern "C" int fvp_precon_overlap(fvp_csr *A, int layers) {
    if (!A) { set_error("null matrix"); return -1; }
    return set_overlap(A, layers);
}

extern "C" int fvp_precon_build(fvp_csr *A, int type) {
    if (!A) { set_error("null matrix"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    return precon_build_impl(A, type);
}

extern "C" int fvp_precon_apply(fvp_csr *A, const double *x, double *y) {
    if (!A || !x || !y) { set_error("null argument"); return -1; }
    if (x == y && A->ptype == FVP_PRECON_ILU0) { set_error("ILU(0) apply cannot run in place"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    return precon_apply_impl(A, x, y);
}

extern "C" int fvp_precon_values(fvp_csr *A, double *values_host) {
    if (!A || !values_host) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    FVP_HIP(hipStreamSynchronize(A->stream));
    if (A->ptype == FVP_PRECON_ILU0)
        FVP_HIP(hipMemcpy(values_host, A->lu.p, sizeof(double) * A->nnz, hipMemcpyDeviceToHost));
    else if (A->ptype == FVP_PRECON_JACOBI)
        FVP_HIP(hipMemcpy(values_host, A->jac.p, sizeof(double) * A->n, hipMemcpyDeviceToHost));
    else if ((A->ptype == FVP_PRECON_BJILU0 || A->ptype == FVP_PRECON_BJILU0_F32) && A->overlap > 0) {
        set_error("overlapping blocks share entries: their factors have no single-pattern form");
        return -1;
    } else if (A->ptype == FVP_PRECON_BJILU0 || A->ptype == FVP_PRECON_BJILU0_F32) {
        // scattered back to the full pattern, 0 off the blocks
        std::vector<double> lu(A->b_lu.n);
        std::vector<int> src(A->b_src.n);
        if (A->ptype == FVP_PRECON_BJILU0)
            FVP_HIP(hipMemcpy(lu.data(), A->b_lu.p, sizeof(double) * lu.size(), hipMemcpyDeviceToHost));
        else {
            std::vector<float> lu32(lu.size());
            FVP_HIP(hipMemcpy(lu32.data(), A->b_lu32.p, sizeof(float) * lu.size(), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < lu.size(); ++k) lu[k] = lu32[k];
        }
        FVP_HIP(hipMemcpy(src.data(), A->b_src.p, sizeof(int) * src.size(), hipMemcpyDeviceToHost));
        std::fill(values_host, values_host + A->nnz, 0.);
        for (size_t k = 0; k < lu.size(); ++k) values_host[src[k]] = lu[k];
    }
    else { set_error("no factored values"); return -1; }
    return 0;
}

extern "C" int fvp_gmres_csr(fvp_csr *A, const double *b, double *x, const fvp_gmres_params *p,
                             fvp_gmres_stats *st) {
    if (!A || !b || !x || !p) { set_error("null argument"); return -1; }
    if (A->ptype < 0) { set_error("preconditioner not built"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    Op op = [A](const double *in, double *out) { return csr_gemv_impl(A, 1.0, in, 0.0, out); };
    Op pc;
    if (A->ptype != FVP_PRECON_NONE) pc = [A](const double *in, double *out) { return precon_apply_impl(A, in, out); };
    return gmres_device(A, A->stream, A->n, op, pc, b, x, p, st);
}

extern "C" int fvp_gmres_csr_deflated(fvp_csr *A, const double *b, double *x, const fvp_gmres_params *p, int k,
                                      fvp_gmres_stats *st) {
    if (!A || !b || !x || !p) { set_error("null argument"); return -1; }
    if (A->ptype < 0) { set_error("preconditioner not built"); return -1; }
    FVP_HIP(hipSetDevice(A->device));
    Op op = [A](const double *in, double *out) { return csr_gemv_impl(A, 1.0, in, 0.0, out); };
    Op pc;
    if (A->ptype != FVP_PRECON_NONE) pc = [A](const double *in, double *out) { return precon_apply_impl(A, in, out); };
    return gmres_deflated(A, A->stream, A->n, op, pc, b, x, p, k, st);
}

// ---- allocation-free GMRES ------------------------------------------------------
// Workspace of fvp_gmres_csr_noalloc: a static device array (32 MiB, one per
// GPU, held until the process ends).  In host code the array's name is only a
// stand-in; its device address on each GPU comes from hipGetSymbolAddress,
// taken once per device with that device current.
constexpr size_t kScratchDoubles = size_t(4) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kScratchDevices];

__global__ void zero_fill(long n, double *__restrict__ p) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.;
}

// the scratch array's address on `device` (which must be current)
static double *scratch_address(int device) {
    if (device < 0 || device >= kScratchDevices) { set_error("device index out of range"); return nullptr; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[device] = static_cast<double *>(ptr);
    }
    return g_scratch_ptr[device];
}

extern "C" int fvp_gmres_csr_noalloc(fvp_csr *A, const double *b, double *x, const fvp_gmres_params *p,
                                     fvp_gmres_stats *st) {
    if (!A || !b || !x || !p) { set_error("null argument"); return -1; }
    if (A->ptype < 0) { set_error("preconditioner not built"); return -1; }
    const int m = p->dim_krylov;
    if (m < 1 || m > 63) { set_error("dim_krylov must be 1..63"); return -1; }
    if (p->reorthogonalise < 0 || p->reorthogonalise > 2) { set_error("reorthogonalise must be 0, 1 or 2"); return -1; }
    const size_t need = gmres_work_doubles(A->n, m);
    if (need > kScratchDoubles) {
        set_error("fvp_gmres_csr_noalloc: the workspace of n = " + std::to_string(A->n) + ", dim_krylov = " +
                  std::to_string(m) + " needs " + std::to_string(need) + " doubles, more than the " +
                  std::to_string(kScratchDoubles) + " of the static workspace (use fvp_gmres_csr)");
        return -1;
    }
    FVP_HIP(hipSetDevice(A->device));
    double *base = scratch_address(A->device);
    if (!base) return -1;
    GmresWork wk = gmres_work_carve(base, A->n, m);
    // the scalars, coefficients and Hessenberg data start cleared on every call
    // (the array keeps the previous call's contents); the basis, work vectors
    // and partials are written before they are read
    const long small = (long)(need - (size_t)(wk.sc - base));
    zero_fill<<<(int)std::min(1024L, (small + 255) / 256), 256, 0, A->stream>>>(small, wk.sc);
    FVP_HIP(hipGetLastError());
    double hstat[2 * 64];
    wk.hstat = hstat;  // pageable: each step's pair is copied and read synchronously
    wk.ev = nullptr;
    wk.hmapped = false;
    Op op = [A](const double *in, double *out) { return csr_gemv_impl(A, 1.0, in, 0.0, out); };
    Op pc;
    if (A->ptype != FVP_PRECON_NONE) pc = [A](const double *in, double *out) { return precon_apply_impl(A, in, out); };
    return gmres_core(wk, A->stream, A->n, op, pc, b, x, p, st);
}

// ================================================================================
// Coloured FD Jacobian and the Newton linear solve on a problem context
// ================================================================================
struct fvp_jacdata {
    fvp_csr *J = nullptr;
    bool coloured = false;             // colouring built (colour_init)
    int ncol = 0;
    std::vector<int> colour;           // per node
    std::vector<int> cptr;             // colour -> [cptr[c], cptr[c+1]) in cols
    std::vector<int> eptr;             // colour -> entries
    DevBuf<int> cols, erow, epos, ecol;
    DevBuf<double> shift, ys, yps, Fs;
    // element-local quotients: the contribution of (element e, row node i,
    // moved node k) goes to srank[16 e + 4 i + k], so that each entry's
    // contributions lie contiguous in [cptr2[p], cptr2[p + 1]) in element
    // order; per row its diagonal entry; k_r and capacity with every node
    // moved by its shift
    DevBuf<long> cptr2;
    DevBuf<int> srank, dpos;
    DevBuf<double> contrib, kr2, cap2;
};

namespace fvp {

__global__ void shift_kernel(int n, double eps, double h, const double *__restrict__ y, const double *__restrict__ yp,
                             const double *__restrict__ ewt, double *__restrict__ shift) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    shift[j] = eps * fmax(fabs(y[j]), fmax(fabs(h * yp[j]), 1.0 / ewt[j]));
}

__global__ void perturb_kernel(int cnt, const int *__restrict__ cols, const double *__restrict__ shift, double c,
                               const double *__restrict__ y, const double *__restrict__ yp, double *__restrict__ ys,
                               double *__restrict__ yps, int restore) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    int j = cols[k];
    if (restore) { ys[j] = y[j]; yps[j] = yp[j]; }
    else { ys[j] = y[j] + shift[j]; yps[j] = yp[j] + c * shift[j]; }
}

// J[pos] = (F0[row] - Fs[row]) / shift[col]   (-(dF/dy + c dF/dy') column entries)
__global__ void fd_scatter(int cnt, const int *__restrict__ erow, const int *__restrict__ epos,
                           const int *__restrict__ ecol, const double *__restrict__ F0,
                           const double *__restrict__ Fs, const double *__restrict__ shift, double *__restrict__ J) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    int r = erow[k];
    J[epos[k]] = (F0[r] - Fs[r]) / shift[ecol[k]];
}

// Node adjacency over element pairs (preconditioner_dss.cpp:107-142) and the
// greedy distance-2 colouring of preconditioner_dss.cpp:72-94.
static int jac_init(fvp_ctx *c) {
    if (c->jd) return 0;
    const int N = c->N, T = c->T, npe = c->npe;
    std::vector<int4> conn(T);
    FVP_HIP(hipMemcpy(conn.data(), c->conn.p, sizeof(int4) * T, hipMemcpyDeviceToHost));
    // node adjacency (the Jacobian pattern): every element's node pairs, counted
    // into per-row segments, then sorted and made unique row by row
    std::vector<int> off(N + 1, 0), buf;
    for (int e = 0; e < T; ++e) {
        const int n[4] = {conn[e].x, conn[e].y, conn[e].z, conn[e].w};
        for (int a = 0; a < npe; ++a) off[n[a] + 1] += npe;
    }
    for (int i = 0; i < N; ++i) off[i + 1] += off[i];
    buf.resize(off[N]);
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int e = 0; e < T; ++e) {
            const int n[4] = {conn[e].x, conn[e].y, conn[e].z, conn[e].w};
            for (int a = 0; a < npe; ++a)
                for (int b = 0; b < npe; ++b) buf[fill[n[a]]++] = n[b];
        }
    }
    std::vector<int> ia(N + 1, 0), ja;
    ja.reserve(off[N] / 4);
    for (int i = 0; i < N; ++i) {
        auto b0 = buf.begin() + off[i], b1 = buf.begin() + off[i + 1];
        std::sort(b0, b1);
        ja.insert(ja.end(), b0, std::unique(b0, b1));
        ia[i + 1] = (int)ja.size();
    }
    std::vector<int>().swap(buf);
    fvp_jacdata *jd = new fvp_jacdata();
    if (csr_create_impl(c->device, N, ia.data(), ja.data(), &jd->J)) { delete jd; return -1; }
    fvp_csr_set_stream(jd->J, c->stream);
    {
        // each entry's contribution slots, counting-sorted by entry (element order within)
        const long nnz = ia[N];
        std::vector<int> pos((size_t)npe * npe * T), dpos(N), srank(16 * (size_t)T, 0);
        std::vector<long> cp(nnz + 1, 0);
        auto at = [&](int row, int col) {
            return (int)(std::lower_bound(ja.begin() + ia[row], ja.begin() + ia[row + 1], col) - ja.begin());
        };
        for (int e = 0; e < T; ++e) {
            const int n[4] = {conn[e].x, conn[e].y, conn[e].z, conn[e].w};
            for (int i = 0; i < npe; ++i)
                for (int k = 0; k < npe; ++k) {
                    const int p = at(n[i], n[k]);
                    pos[((size_t)e * npe + i) * npe + k] = p;
                    ++cp[p + 1];
                }
        }
        for (long p = 0; p < nnz; ++p) cp[p + 1] += cp[p];
        {
            std::vector<long> fill(cp.begin(), cp.end() - 1);
            for (int e = 0; e < T; ++e)
                for (int i = 0; i < npe; ++i)
                    for (int k = 0; k < npe; ++k)
                        srank[16 * (size_t)e + 4 * i + k] = (int)fill[pos[((size_t)e * npe + i) * npe + k]]++;
        }
        for (int i = 0; i < N; ++i) dpos[i] = at(i, i);
        if ((long)16 * T > INT32_MAX) { set_error("too many elements for the FD Jacobian slots"); fvp_csr_destroy(jd->J); delete jd; return -1; }
        if (jd->cptr2.upload(cp.data(), cp.size()) || jd->srank.upload(srank.data(), srank.size()) ||
            jd->dpos.upload(dpos.data(), N) || jd->contrib.alloc((size_t)npe * npe * T) ||
            jd->kr2.alloc((size_t)c->nz * N) || jd->cap2.alloc(N)) {
            fvp_csr_destroy(jd->J);
            delete jd;
            return -1;
        }
    }
    if (jd->shift.alloc(N) || jd->ys.alloc(N) || jd->yps.alloc(N) || jd->Fs.alloc(N)) {
        fvp_csr_destroy(jd->J);
        delete jd;
        return -1;
    }
    c->jd = jd;
    return 0;
}

// The reference's greedy distance-2 colouring in node order and, per colour,
// the entries its columns fill (the coloured FD Jacobian and fvp_colours only;
// built on first use)
static int colour_init(fvp_ctx *c) {
    if (jac_init(c)) return -1;
    fvp_jacdata *jd = c->jd;
    if (jd->coloured) return 0;
    const int N = c->N;
    const std::vector<int> &ia = jd->J->ia_h, &ja = jd->J->ja_h;
    // greedy distance-2 colouring in node order
    std::vector<int> colour(N, N - 1), mark(N, -1);
    int maxc = 0;
    for (int i = 0; i < N; ++i) {
        for (int p = ia[i]; p < ia[i + 1]; ++p) {
            int nb = ja[p];
            for (int q = ia[nb]; q < ia[nb + 1]; ++q) mark[colour[ja[q]]] = i;
        }
        int sc = 0;
        while (sc < maxc && mark[sc] == i) ++sc;
        if (sc == maxc) ++maxc;
        colour[i] = sc;
    }
    jd->ncol = maxc;
    jd->colour = colour;
    std::vector<int> cols(N), cptr(maxc + 1, 0);
    for (int i = 0; i < N; ++i) cptr[colour[i] + 1]++;
    for (int k = 0; k < maxc; ++k) cptr[k + 1] += cptr[k];
    {
        std::vector<int> fill(cptr.begin(), cptr.end() - 1);
        for (int i = 0; i < N; ++i) cols[fill[colour[i]]++] = i;
    }
    // entries per colour: for each column j of the colour, rows = pattern of column j
    std::vector<int> erow, epos, ecol, eptr(maxc + 1, 0);
    erow.reserve(ia[N]); epos.reserve(ia[N]); ecol.reserve(ia[N]);
    for (int k = 0; k < maxc; ++k) {
        for (int q = cptr[k]; q < cptr[k + 1]; ++q) {
            int j = cols[q];
            for (int p = ia[j]; p < ia[j + 1]; ++p) {
                int row = ja[p];
                // position of (row, j) in row's sorted column list
                int lo = ia[row], hi = ia[row + 1];
                int pos = (int)(std::lower_bound(ja.begin() + lo, ja.begin() + hi, j) - ja.begin());
                erow.push_back(row); epos.push_back(pos); ecol.push_back(j);
            }
        }
        eptr[k + 1] = (int)erow.size();
    }
    jd->cptr = cptr;
    jd->eptr = eptr;
    if (jd->cols.upload(cols.data(), N) || jd->erow.upload(erow.data(), erow.size()) ||
        jd->epos.upload(epos.data(), epos.size()) || jd->ecol.upload(ecol.data(), ecol.size()))
        return -1;
    jd->coloured = true;
    return 0;
}

fvp_csr *jacobian_handle(fvp_ctx *c) {
    if (jac_init(c)) return nullptr;
    fvp_csr_set_stream(c->jd->J, c->stream);
    return c->jd->J;
}

void jac_free(fvp_ctx *c) {
    if (!c->jd) return;
    fvp_jacdata *jd = c->jd;
    fvp_csr_destroy(jd->J);
    jd->cols.release(); jd->erow.release(); jd->epos.release(); jd->ecol.release();
    jd->shift.release(); jd->ys.release(); jd->yps.release(); jd->Fs.release();
    jd->cptr2.release(); jd->srank.release(); jd->dpos.release(); jd->contrib.release();
    jd->kr2.release(); jd->cap2.release();
    delete jd;
    c->jd = nullptr;
}

}  // namespace fvp

extern "C" int fvp_precon_setup(fvp_ctx *c, double t, double c_j, double h, const double *psi, const double *psip,
                                const double *F0, const double *ewt, int type) {
    if (!c || !psi || !psip || !F0 || !ewt) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    if (jac_init(c)) return -1;
    fvp_jacdata *jd = c->jd;
    hipStream_t s = c->stream;
    fvp_csr_set_stream(jd->J, s);
    const int N = c->N;
    const double eps = 1.4901161193847656e-08;  // sqrt(2^-52)
    shift_kernel<<<nblk(N, 256), 256, 0, s>>>(N, eps, h, psi, psip, ewt, jd->shift.p);
    if (!c->averaging && c->jac_mode == FVP_JACOBIAN_ELEMENT) {
        if (fd_jacobian_elem(c, c_j, psi, psip, jd->shift.p, jd->ys.p, jd->kr2.p, jd->cap2.p, jd->contrib.p,
                             jd->cptr2.p, jd->srank.p, jd->J->ia.p, jd->dpos.p, jd->J->a.p, jd->J->nnz,
                             type == FVP_PRECON_BJILU0_F32))
            return -1;
        return precon_build_impl(jd->J, type);
    }
    if (colour_init(c)) return -1;
    vcopy<<<nblk(N, 256), 256, 0, s>>>(N, psi, jd->ys.p);
    vcopy<<<nblk(N, 256), 256, 0, s>>>(N, psip, jd->yps.p);
    for (int k = 0; k < jd->ncol; ++k) {
        int nc = jd->cptr[k + 1] - jd->cptr[k];
        int ne = jd->eptr[k + 1] - jd->eptr[k];
        perturb_kernel<<<nblk(nc, 256), 256, 0, s>>>(nc, jd->cols.p + jd->cptr[k], jd->shift.p, c_j, psi, psip,
                                                     jd->ys.p, jd->yps.p, 0);
        if (residual_impl(c, t, jd->ys.p, jd->yps.p, jd->Fs.p, c->averaging)) return -1;
        fd_scatter<<<nblk(ne, 256), 256, 0, s>>>(ne, jd->erow.p + jd->eptr[k], jd->epos.p + jd->eptr[k],
                                                 jd->ecol.p + jd->eptr[k], F0, jd->Fs.p, jd->shift.p, jd->J->a.p);
        perturb_kernel<<<nblk(nc, 256), 256, 0, s>>>(nc, jd->cols.p + jd->cptr[k], jd->shift.p, c_j, psi, psip,
                                                     jd->ys.p, jd->yps.p, 1);
    }
    FVP_HIP(hipGetLastError());
    return precon_build_impl(jd->J, type);
}

namespace fvp {
// z = M^-1 (r / rdiv): the integrator's scaled solve (tmp: a work vector of N)
int precon_solve_div(fvp_ctx *c, const double *r, const double *rdiv, double *tmp, double *z, const double **zn2,
                     int *zn2_count) {
    if (!c->jd || c->jd->J->ptype < 0) { set_error("fvp_precon_setup has not run"); return -1; }
    fvp_csr *A = c->jd->J;
    fvp_csr_set_stream(A, c->stream);
    const bool bj = A->ptype == FVP_PRECON_BJILU0 || A->ptype == FVP_PRECON_BJILU0_F32;
    double *part = nullptr;
    if (zn2 && bj) {
        if ((size_t)A->nblk > A->b_zn2.n && A->b_zn2.alloc(A->nblk)) return -1;
        part = A->b_zn2.p;
    }
    prof_mark(c, 2, 0);
    int rc = precon_apply_impl(A, r, z, rdiv, tmp, part);
    prof_mark(c, 2, 1);
    if (zn2) *zn2 = part;
    if (zn2_count) *zn2_count = part ? A->nblk : 0;
    return rc;
}
}  // namespace fvp

extern "C" int fvp_precon_solve(fvp_ctx *c, const double *r, double *z) {
    if (!c || !r || !z) { set_error("null argument"); return -1; }
    if (!c->jd || c->jd->J->ptype < 0) { set_error("fvp_precon_setup has not run"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    fvp_csr_set_stream(c->jd->J, c->stream);
    prof_mark(c, 2, 0);
    int rc = precon_apply_impl(c->jd->J, r, z);
    prof_mark(c, 2, 1);
    return rc;
}

extern "C" int fvp_set_jacobian_mode(fvp_ctx *c, int mode) {
    if (!c) { set_error("null argument"); return -1; }
    if (mode != FVP_JACOBIAN_ELEMENT && mode != FVP_JACOBIAN_COLOURED) { set_error("invalid Jacobian mode"); return -1; }
    c->jac_mode = mode;
    return 0;
}

extern "C" int fvp_jacobian(fvp_ctx *c, fvp_csr **A) {
    if (!c || !A) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    if (jac_init(c)) return -1;
    *A = c->jd->J;
    return 0;
}

extern "C" int fvp_set_block_rows(fvp_ctx *c, int rows) {
    if (!c) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    if (jac_init(c)) return -1;
    return set_block_rows(c->jd->J, rows);
}

extern "C" int fvp_set_overlap(fvp_ctx *c, int layers) {
    if (!c) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    if (jac_init(c)) return -1;
    return set_overlap(c->jd->J, layers);
}

extern "C" int fvp_colours(fvp_ctx *c, int *n_colours, int *colour_host) {
    if (!c || !n_colours) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    if (colour_init(c)) return -1;
    *n_colours = c->jd->ncol;
    if (colour_host) std::copy(c->jd->colour.begin(), c->jd->colour.end(), colour_host);
    return 0;
}

extern "C" int fvp_newton_gmres(fvp_ctx *c, double t, const double *psi, const double *psip, const double *F0,
                                double c_j, const double *b, double *x, const fvp_gmres_params *p,
                                fvp_gmres_stats *st) {
    if (!c || !psi || !psip || !F0 || !b || !x || !p) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    if (jac_init(c)) return -1;
    fvp_csr *W = c->jd->J;
    fvp_csr_set_stream(W, c->stream);
    Op op = [=](const double *in, double *out) { return fvp_jv(c, t, psi, psip, F0, c_j, in, out); };
    Op pc;
    if (W->ptype > 0) pc = [W](const double *in, double *out) { return precon_apply_impl(W, in, out); };
    return gmres_device(W, c->stream, c->N, op, pc, b, x, p, st);
}
