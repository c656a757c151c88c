// Small pointwise kernels outside the split step: the spin-density observable and the
// packing of real fields for the spectral-derivative entry points.
#include <hip/hip_runtime.h>

#include "bec_aux.h"
#include "bec_device.h"
#include "bec_kernels.h"

namespace bec {

// observables.jl:61-88 (_compute_spin_density!): per site, in FP64,
//   Fz = sum_c m_c |psi_c|^2,   Fx + i Fy = sum_{c >= 1} f+(m_c) conj(psi_{c-1}) psi_c,
// f+(m) = sqrt(F(F+1) - m(m+1)), m_c = F - c.  out: [3][nsites] = Fx, Fy, Fz.
template <typename T>
__global__ void spin_density_kernel(const cplx<T>* __restrict__ psi, double* __restrict__ out,
                                    long long nsites, int D) {
  const int F = (D - 1) / 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nsites;
       i += (long long)gridDim.x * blockDim.x) {
    double fx = 0.0, fy = 0.0, fz = 0.0;
    cplx<T> prev = psi[i];
    fz = (double)F * ((double)prev.x * prev.x + (double)prev.y * prev.y);
    for (int c = 1; c < D; ++c) {
      const cplx<T> cur = psi[(long long)c * nsites + i];
      const double m = (double)(F - c);
      fz += m * ((double)cur.x * cur.x + (double)cur.y * cur.y);
      const double fp = sqrt((double)(F * (F + 1)) - m * (m + 1.0));
      // conj(prev) * cur
      fx += fp * ((double)prev.x * cur.x + (double)prev.y * cur.y);
      fy += fp * ((double)prev.x * cur.y - (double)prev.y * cur.x);
      prev = cur;
    }
    out[i] = fx;
    out[nsites + i] = fy;
    out[2 * nsites + i] = fz;
  }
}

template <typename T>
hipError_t launch_spin_density(hipStream_t st, const void* psi, double* out, long long nsites,
                               int D) {
  long long nb = (nsites + 255) / 256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL((spin_density_kernel<T>), dim3((unsigned)nb), dim3(256), 0, st,
                     (const cplx<T>*)psi, out, nsites, D);
  return hipGetLastError();
}

__global__ void grad_pack_kernel(const double* __restrict__ f, cplx<double>* __restrict__ psi,
                                 long long n, int nd) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double v = f[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) psi[d * n + i] = mk(d < nd ? v : 0.0, 0.0);
  }
}

__global__ void grad_unpack_kernel(const cplx<double>* __restrict__ psi, double* __restrict__ out,
                                   long long n, int nd) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    for (int d = 0; d < nd; ++d) out[d * n + i] = psi[d * n + i].x;
}

static unsigned grid_of(long long n) {
  long long nb = (n + 255) / 256;
  return (unsigned)(nb > 8192 ? 8192 : (nb < 1 ? 1 : nb));
}

hipError_t launch_grad_pack(hipStream_t st, const double* f, void* psi, long long n, int nd) {
  hipLaunchKernelGGL(grad_pack_kernel, dim3(grid_of(n)), dim3(256), 0, st, f,
                     (cplx<double>*)psi, n, nd);
  return hipGetLastError();
}

hipError_t launch_grad_unpack(hipStream_t st, const void* psi, double* out, long long n, int nd) {
  hipLaunchKernelGGL(grad_unpack_kernel, dim3(grid_of(n)), dim3(256), 0, st,
                     (const cplx<double>*)psi, out, n, nd);
  return hipGetLastError();
}

template hipError_t launch_spin_density<float>(hipStream_t, const void*, double*, long long, int);
template hipError_t launch_spin_density<double>(hipStream_t, const void*, double*, long long, int);

}  // namespace bec
