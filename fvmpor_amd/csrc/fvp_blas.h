// Device vector helpers and the GMRES driver shared by the Krylov, preconditioner
// and integrator code.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

#include "../../include/fvp.h"

namespace fvp {
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block sum of v (thread 0 holds the result; lds: one double per wavefront)
__device__ __forceinline__ double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += lds[w];
    __syncthreads();
    return r;
}

// FD increment of J v: sigma = sqrt(eps) ||psi|| / ||v|| from the squared norms
__host__ __device__ inline double jv_sigma_of(double psi_n2, double v_n2) {
    const double sq_eps = 1.4901161193847656e-08;  // sqrt(2^-52)
    double np_ = sqrt(psi_n2), nv = sqrt(v_n2);
    if (np_ == 0.) np_ = 1.;
    return nv > 0. ? sq_eps * np_ / nv : 1.;
}

// out[0] (device) = x . y; part must hold 512 doubles.  Asynchronous on s.
// hout (optional): pinned host memory that receives the value as well.
int dot_async(hipStream_t s, long n, const double *x, const double *y, double *part, double *out,
              double *hout = nullptr);
// Waits until the work queued on s so far has finished: records ev behind it
// and polls the event.  Scalars that kernels stored to pinned host memory are
// then readable; the poll avoids the wake-up latency of a blocking wait.
int host_wait(hipStream_t s, hipEvent_t ev);

using Op = std::function<int(const double *, double *)>;
// Restarted right-preconditioned GMRES on A x = b; W supplies the workspace.
// x_zero: the caller has set x = 0, so the first residual is b itself (no
// product with A).
int gmres_device(fvp_csr *W, hipStream_t s, long n, const Op &A, const Op &M, const double *b, double *x,
                 const fvp_gmres_params *p, fvp_gmres_stats *st, bool x_zero = false);
// the context's Jacobian handle (built on first use)
fvp_csr *jacobian_handle(fvp_ctx *c);
}  // namespace fvp
