// y / z passes for gfx950: tiles of TB contiguous x columns by the whole line length,
// Stockham FFT in LDS; the last axis does forward FFT -> k-space multiply -> inverse FFT in
// the same launch (kinetic phase, or C_dd Q_ab(k) for the DDI spectra).  Also the Q_ab
// product of 1-D DDI spectra, which have no transverse pass to carry it.
// (A translation unit of its own so that it compiles beside the x pass.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "bec_device.h"
#include "bec_kernels.h"
#include "bec_passes.h"

namespace bec {

// ---------------------------------------------------------------------------
// y / z pass
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE, int TB, int FIELDS>
__global__ void __launch_bounds__(1024)
strided_kernel(Geom g, DevTables tb, cplx<T>* __restrict__ data, int axis, double kin_frac,
               double dt_override, int itp, int ddi_data, int order) {
  // Block order (the dispatch order of the tiles): bit 0 reverses the tiles of a component
  // (for the y pass: the z planes from the last one), bit 1 the components, so that a pass
  // starts with the data its predecessor wrote last, which the L2 and the memory-side cache
  // still hold (psi is larger than the 256 MB memory-side cache).
  const int bx_ = (order & 1) ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x;
  const int by_ = (order & 2) ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;
#define BEC_BX bx_
#define BEC_BY by_
#define BEC_BZ blockIdx.z
#define BEC_GX gridDim.x
#include "bec_strided_body.h"
#undef BEC_BX
#undef BEC_BY
#undef BEC_BZ
#undef BEC_GX
}

int strided_tb(int N, int fields, int csize) {
  return fft_size_id(N) >= 0 ? strided_tb_c(N, fields, csize) : sgen_tb(N, fields, csize);
}

// Lines of the other transverse axis that a y/z pass visits.  Zero-padded DDI spectra: the
// y passes only touch the z < n[2] planes that hold data (or are kept); the z pass all sn[1].
static int strided_other(const Geom& g, int axis, int ddi_data) {
  if (!ddi_data) return axis == 1 ? g.n[2] : g.n[1];
  return axis == 1 ? g.n[2] : g.sn[1];
}

// Block order of a psi pass (bit 0: tiles of a component from the last, bit 1: components
// from the last).  A 3-D kinetic step is x pass -> y FWD -> z KIN -> y INV -> x pass, and
// psi (436 MB at 128^3 spin-6 FP64) is larger than the memory-side cache: the y passes
// visit the components in order, so the z pass takes them from the last one (the data the
// y pass wrote last), and the y inverse pass, in order again, starts with what the z pass
// wrote last.  Measured at 128^3 spin-6 FP64: z pass 186 -> 164 us, y inverse 171 -> 135 us;
// reversing the z planes of the y forward pass as well was 5 % slower.
// BEC_STRIDED_ORDER=0 keeps the plain order everywhere (a measurement switch).
static int strided_order(const Geom& g, int axis, int mode, int ddi_data) {
  static const int enabled = [] {
    const char* e = getenv("BEC_STRIDED_ORDER");
    return e ? atoi(e) : 1;
  }();
  (void)mode;
  if (!enabled || ddi_data || g.ndim != 3) return 0;
  return axis == 2 ? 2 : 0;
}

template <typename T, int N, int MODE, int FIELDS>
static hipError_t strided_n(hipStream_t st, const Geom& g, const DevTables& tb, int axis,
                            void* data, int n_active, double kin_frac, double dt_override,
                            int itp, int ddi_data) {
  // TB columns per tile: keep the tile at <= 64 KiB of LDS
  constexpr int TB = strided_tb_c(N, FIELDS, (int)sizeof(cplx<T>));
  const int valid = ddi_data ? g.nxh : g.n[0];
  const int nxch = (valid + TB - 1) / TB;
  const int n_other = strided_other(g, axis, ddi_data);
  const int ngroups = ddi_data ? 1 : g.D / FIELDS;
  dim3 grid(nxch * n_other, ngroups, n_active);
  const int threads = FIELDS * TB * (N / fft_e<N>());
  const size_t lds = sizeof(cplx<T>) * FIELDS * N * TB + sizeof(double) * (2 * (N + TB + 1) + 16) +
                     sizeof(cplx<double>) * N;  // twiddles
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)strided_kernel<T, N, MODE, TB, FIELDS>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((strided_kernel<T, N, MODE, TB, FIELDS>), grid, dim3(threads), lds, st, g, tb,
                     (cplx<T>*)data, axis, kin_frac, dt_override, itp, ddi_data,
                     strided_order(g, axis, MODE, ddi_data));
  return hipGetLastError();
}

template <typename T, int N>
static hipError_t strided_mode(hipStream_t st, const Geom& g, const DevTables& tb, int axis,
                               int mode, void* data, int n_active, double kin_frac,
                               double dt_override, int itp, int ddi_data) {
  if (ddi_data) {
    switch (mode) {
      case SM_FWD: return strided_n<T, N, SM_FWD, 3>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 1);
      case SM_INV: return strided_n<T, N, SM_INV, 3>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 1);
      case SM_DDI: return strided_n<T, N, SM_DDI, 3>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 1);
      default: return hipErrorInvalidValue;
    }
  }
  switch (mode) {
    case SM_FWD: return strided_n<T, N, SM_FWD, 1>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 0);
    case SM_INV: return strided_n<T, N, SM_INV, 1>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 0);
    case SM_KIN: return strided_n<T, N, SM_KIN, 1>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 0);
    case SM_KE: return strided_n<T, N, SM_KE, 1>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 0);
    case SM_DERIV:
      if constexpr (sizeof(T) == 8)
        return strided_n<T, N, SM_DERIV, 1>(st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, 0);
      else
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_strided(hipStream_t st, const Geom& g, const DevTables& tb, int axis, int mode,
                          void* data, int n_active, double kin_frac, double dt_override, int itp,
                          int ddi_data, int* nblocks_out) {
  const int len = ddi_data ? g.sn[axis] : g.n[axis];
  if (nblocks_out) {
    const int tbw = strided_tb(len, ddi_data ? 3 : 1, (int)sizeof(cplx<T>));
    const int valid = ddi_data ? g.nxh : g.n[0];
    const int n_other = strided_other(g, axis, ddi_data);
    *nblocks_out = ((valid + tbw - 1) / tbw) * n_other * (ddi_data ? 1 : g.D);
  }
  switch (len) {
    case 8: return strided_mode<T, 8>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 24: return strided_mode<T, 24>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 40: return strided_mode<T, 40>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 48: return strided_mode<T, 48>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 80: return strided_mode<T, 80>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 96: return strided_mode<T, 96>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 160: return strided_mode<T, 160>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 192: return strided_mode<T, 192>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 56: return strided_mode<T, 56>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 112: return strided_mode<T, 112>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 224: return strided_mode<T, 224>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 320: return strided_mode<T, 320>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 384: return strided_mode<T, 384>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 16: return strided_mode<T, 16>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 32: return strided_mode<T, 32>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 64: return strided_mode<T, 64>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 128: return strided_mode<T, 128>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 256: return strided_mode<T, 256>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    case 512: return strided_mode<T, 512>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override, itp, ddi_data);
    default:  // no built FFT of this length
      return launch_strided_gen<T>(st, g, tb, axis, mode, data, n_active, kin_frac, dt_override,
                                   itp, ddi_data, strided_other(g, axis, ddi_data));
  }
}

// 1-D grids: the x pass leaves the three half spectra (nxh bins each) and no transverse
// pass follows to carry the k-space product, so it runs on its own, one bin per thread
// (ddi.jl:108-136 with N = 1: k = k_x, rfftfreq with the Nyquist bin positive).
template <typename T>
__global__ void __launch_bounds__(64) ddi_line_kernel(Geom g, DevTables tb, cplx<T>* __restrict__ fhat) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.nxh) return;
  const int inst = tb.active[blockIdx.y];
  const double kx = (x == g.sn[0] / 2) ? -tb.sktab[g.sn[0] / 2] : tb.sktab[x];
  const double s = tb.params[inst].cdd / (double)g.sn[0];
  cplx<T>* f = fhat + (size_t)inst * g.fhat_inst_stride + x;
  cplx<T> fx = f[0], fy = f[g.fhat_field_stride], fz = f[2 * g.fhat_field_stride];
  ddi_qmix<T>(kx, 0.0, 0.0, 1, g.ddi_secular != 0, s, fx, fy, fz);
  f[0] = fx;
  f[g.fhat_field_stride] = fy;
  f[2 * g.fhat_field_stride] = fz;
}

template <typename T>
hipError_t launch_ddi_line(hipStream_t st, const Geom& g, const DevTables& tb, void* fhat,
                           int n_active) {
  if (g.ndim != 1) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((g.nxh + 63) / 64), n_active);
  hipLaunchKernelGGL((ddi_line_kernel<T>), grid, dim3(64), 0, st, g, tb, (cplx<T>*)fhat);
  return hipGetLastError();
}
template hipError_t launch_ddi_line<float>(hipStream_t, const Geom&, const DevTables&, void*, int);
template hipError_t launch_ddi_line<double>(hipStream_t, const Geom&, const DevTables&, void*, int);

template hipError_t launch_strided<float>(hipStream_t, const Geom&, const DevTables&, int, int,
                                            void*, int, double, double, int, int, int*);
template hipError_t launch_strided<double>(hipStream_t, const Geom&, const DevTables&, int, int,
                                             void*, int, double, double, int, int, int*);

}  // namespace bec
