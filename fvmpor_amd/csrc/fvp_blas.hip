// Vector kernels of the Newton-Krylov loop: wavefront-shuffle reductions
// (norms, fused multi-dots for Gram-Schmidt), axpy forms, and the
// finite-difference Jacobian-vector product.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fvp_blas.h"
#include "fvp_internal.h"

namespace fvp {

constexpr int kRedBlocks = 512;   // partial sums per reduction (>= 2 per CU)
constexpr int kRedThreads = 256;

// partials[b] = sum over this block's grid-stride slice of x*y (x == y gives ||x||^2)
__global__ void __launch_bounds__(kRedThreads) dot_partial(long n, const double *__restrict__ x,
                                                          const double *__restrict__ y, double *__restrict__ part) {
    __shared__ double lds[kRedThreads / 64];
    double s = 0.;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += x[i] * y[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[0] = sum of np partials (fixed order: deterministic); hout (optional,
// pinned host memory) gets it too
__global__ void __launch_bounds__(kRedThreads) sum_partials(int np, const double *__restrict__ part,
                                                           double *__restrict__ out, double *__restrict__ hout) {
    __shared__ double lds[kRedThreads / 64];
    // each lane's partials in order, four loads in flight at a time
    double s = 0.;
    const int bd = blockDim.x;
    for (int i0 = threadIdx.x; i0 < np; i0 += 4 * bd) {
        double q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = i0 + u * bd < np ? part[i0 + u * bd] : 0.;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u * bd < np) s += q[u];
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        out[0] = s;
        if (hout) {
            hout[0] = s;
            __threadfence_system();
        }
    }
}

// FD increment sigma = sqrt(eps) * ||psi|| / ||v|| from two squared norms
__global__ void __launch_bounds__(kRedThreads) jv_sigma(int np, const double *__restrict__ part_psi,
                                                       const double *__restrict__ part_v, double *__restrict__ sig) {
    __shared__ double lds[kRedThreads / 64];
    double a = 0., b = 0.;
    for (int i = threadIdx.x; i < np; i += blockDim.x) { a += part_psi[i]; b += part_v[i]; }
    a = block_sum(a, lds);
    b = block_sum(b, lds);
    if (threadIdx.x == 0) sig[0] = jv_sigma_of(a, b);
}

// the same from ||psi||^2 already reduced (psi_n2[0]) and the partials of ||v||^2
__global__ void __launch_bounds__(kRedThreads) jv_sigma_n2(int np, const double *__restrict__ psi_n2,
                                                          const double *__restrict__ part_v, double *__restrict__ sig) {
    __shared__ double lds[kRedThreads / 64];
    double b = 0.;
    for (int i = threadIdx.x; i < np; i += blockDim.x) b += part_v[i];
    b = block_sum(b, lds);
    if (threadIdx.x == 0) sig[0] = jv_sigma_of(psi_n2[0], b);
}

__global__ void jv_perturb(long n, const double *__restrict__ sig, double cj, const double *__restrict__ psi,
                           const double *__restrict__ psip, const double *__restrict__ v, double *__restrict__ y,
                           double *__restrict__ yp) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = sig[0];
    y[i] = psi[i] + s * v[i];
    yp[i] = psip[i] + cj * (s * v[i]);
}

__global__ void jv_finish(long n, const double *__restrict__ sig, const double *__restrict__ F1,
                          const double *__restrict__ F0, const double *__restrict__ osc, double *__restrict__ Jv) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = (F1[i] - F0[i]) / sig[0];
    Jv[i] = osc ? d * osc[i] : d;
}

static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

int dot_async(hipStream_t s, long n, const double *x, const double *y, double *part, double *out, double *hout) {
    dot_partial<<<kRedBlocks, kRedThreads, 0, s>>>(n, x, y, part);
    sum_partials<<<1, kRedThreads, 0, s>>>(kRedBlocks, part, out, hout);
    FVP_HIP(hipGetLastError());
    return 0;
}

int host_wait(hipStream_t s, hipEvent_t ev) {
    FVP_HIP(hipEventRecord(ev, s));
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
    }
    FVP_HIP(e);
    return 0;
}

}  // namespace fvp

using namespace fvp;

namespace fvp {
int jv_impl(fvp_ctx *c, double t, const double *psi, const double *psip, const double *F0, double c_j,
            const double *v, double *Jv, const double *psi_n2, const double *osc, const double *v_n2,
            int v_count) {
    hipStream_t s = c->stream;
    const long N = c->N;
    double *part = c->red.p;  // [0, 512) psi partials, [512, 1024) v partials, [1024] sigma
    // partials of ||v||^2 from the solve that wrote v: the fused residual
    // kernels' van Genuchten pass forms sigma itself
    const bool fold = psi_n2 && v_n2 && !c->averaging;
    if (!fold && psi_n2 && v_n2) {
        jv_sigma_n2<<<1, kRedThreads, 0, s>>>(v_count, psi_n2, v_n2, part + 2 * kRedBlocks);
    } else if (!fold && psi_n2) {
        dot_partial<<<kRedBlocks, kRedThreads, 0, s>>>(N, v, v, part + kRedBlocks);
        jv_sigma_n2<<<1, kRedThreads, 0, s>>>(kRedBlocks, psi_n2, part + kRedBlocks, part + 2 * kRedBlocks);
    } else if (!fold) {
        dot_partial<<<kRedBlocks, kRedThreads, 0, s>>>(N, v, v, part + kRedBlocks);
        dot_partial<<<kRedBlocks, kRedThreads, 0, s>>>(N, psi, psi, part);
        jv_sigma<<<1, kRedThreads, 0, s>>>(kRedBlocks, part, part + kRedBlocks, part + 2 * kRedBlocks);
    }
    FVP_HIP(hipGetLastError());
    if (!c->averaging) {  // perturbation and difference quotient fused into the residual kernels
        JvArgs jv;
        jv.v = v; jv.sig = part + 2 * kRedBlocks; jv.F0 = F0; jv.cj = c_j; jv.osc = osc;
        if (fold) {
            jv.psi_n2 = psi_n2; jv.part_v = v_n2; jv.np = v_count; jv.sig_out = part + 2 * kRedBlocks;
        }
        return residual_impl(c, t, psi, psip, Jv, 0, &jv);
    }
    jv_perturb<<<nblk(N, 256), 256, 0, s>>>(N, part + 2 * kRedBlocks, c_j, psi, psip, v, c->tmp1.p, c->tmp2.p);
    FVP_HIP(hipGetLastError());
    if (residual_impl(c, t, c->tmp1.p, c->tmp2.p, c->tmp3.p, c->averaging)) return -1;
    jv_finish<<<nblk(N, 256), 256, 0, s>>>(N, part + 2 * kRedBlocks, c->tmp3.p, F0, osc, Jv);
    FVP_HIP(hipGetLastError());
    return 0;
}
}  // namespace fvp

extern "C" int fvp_jv(fvp_ctx *c, double t, const double *psi, const double *psip, const double *F0, double c_j,
                      const double *v, double *Jv) {
    if (!c || !psi || !psip || !F0 || !v || !Jv) { set_error("null argument"); return -1; }
    FVP_HIP(hipSetDevice(c->device));
    return jv_impl(c, t, psi, psip, F0, c_j, v, Jv, nullptr, nullptr);
}
