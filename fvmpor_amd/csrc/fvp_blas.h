// Device vector helpers shared by the Krylov and preconditioner code.
#pragma once
#include <hip/hip_runtime.h>

namespace fvp {
// out[0] (device) = x . y; part must hold 512 doubles.  Asynchronous on s.
int dot_async(hipStream_t s, long n, const double *x, const double *y, double *part, double *out);
}  // namespace fvp
