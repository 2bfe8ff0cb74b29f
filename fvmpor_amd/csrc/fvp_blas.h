// Device vector helpers and the GMRES driver shared by the Krylov, preconditioner
// and integrator code.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

#include "../../include/fvp.h"

namespace fvp {
// out[0] (device) = x . y; part must hold 512 doubles.  Asynchronous on s.
int dot_async(hipStream_t s, long n, const double *x, const double *y, double *part, double *out);

using Op = std::function<int(const double *, double *)>;
// Restarted right-preconditioned GMRES on A x = b; W supplies the workspace.
int gmres_device(fvp_csr *W, hipStream_t s, long n, const Op &A, const Op &M, const double *b, double *x,
                 const fvp_gmres_params *p, fvp_gmres_stats *st);
// the context's Jacobian handle (built on first use)
fvp_csr *jacobian_handle(fvp_ctx *c);
}  // namespace fvp
