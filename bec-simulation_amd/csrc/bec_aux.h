// Launchers of the small pointwise kernels in bec_aux.hip (used by bec_capi.hip only).
#pragma once
#include <hip/hip_runtime.h>

namespace bec {

// Real field f[N] into the complex components of a spin-1 psi [3][N]: component d < nd
// holds (f, 0), the others (0, 0).
hipError_t launch_grad_pack(hipStream_t st, const double* f, void* psi, long long n, int nd);
// out[d][i] = Re psi[d][i] for d < nd.
hipError_t launch_grad_unpack(hipStream_t st, const void* psi, double* out, long long n, int nd);

}  // namespace bec
