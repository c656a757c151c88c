// Grouped launches: one kernel launch per schedule position for several workspaces of
// different grid shapes (the batched ground-state solves of bec_group_split_step).
//
// Every workspace of a group runs the same schedule (same dimension, spin, precision and
// operator set), so position p of every workspace's launch sequence is the same kind of
// pass.  The grouped kernel of that pass takes a device array of descriptors (geometry,
// tables, program, buffers, and the linear index of the descriptor's first block); a block
// finds its descriptor from its own index, and runs the tile of that workspace with the
// same code as the single-workspace kernel (the kernel bodies are pasted from
// bec_xpass_body.h and bec_strided_body.h).  The x pass keeps each workspace's own tiles
// (so its per-block partial sums are those of the workspace's own launches); the y/z passes
// use 2048-element tiles for every line length (2048 / N columns), so that one block size
// serves all shapes, and compute every column exactly as the single-workspace pass does.
//
// Compiled once per part: BEC_PART = 0 holds the y/z kernels and the dispatch, BEC_PART = D
// (1, 3, 5) the x-pass kernel of D components.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bec_device.h"
#include "bec_kernels.h"
#include "bec_passes.h"

#ifndef BEC_PART
#error "compile with -DBEC_PART=0 (y/z kernels, dispatch) or -DBEC_PART=D (x pass of D components)"
#endif

namespace bec {

template <typename Desc>
__device__ __forceinline__ const Desc& group_desc(const Desc* d, int nd) {
  int k = 0;
  for (int i = 1; i < nd; ++i)
    if ((int)blockIdx.x >= d[i].block0) k = i;
  return d[k];
}

template <typename T, int D>
hipError_t xgroup_launch_d(hipStream_t st, const XGroupDesc* descs, int nd, int blocks,
                           int threads, size_t lds);

#if BEC_PART != 0
// ---------------------------------------------------------------------------
// x pass
// ---------------------------------------------------------------------------
template <typename T, int NX, int D>
__device__ __forceinline__ void xpass_group_tile(const XGroupDesc& x, int bx_, int by_) {
  const Geom& g = x.g;
  const DevTables& tb = x.tb;
  const XProgram& prog = x.prog;
  const cplx<T>* __restrict__ src = static_cast<const cplx<T>*>(x.src);
  cplx<T>* __restrict__ dst = static_cast<cplx<T>*>(x.dst);
  cplx<T>* __restrict__ fhat = static_cast<cplx<T>*>(x.fhat);
  T* __restrict__ phi = static_cast<T*>(x.phi);
  const int ROWS = x.rows;
#define BEC_BX bx_
#define BEC_BY by_
#define BEC_REPEAT 0
#include "bec_xpass_body.h"
#undef BEC_REPEAT
#undef BEC_BX
#undef BEC_BY
}

// The x-pass descriptors travel as the kernel's argument and are read in place through the
// kernel-argument segment (constant memory, which the code generator may re-read at will,
// as it does the single-workspace kernel's arguments): descriptors in global memory, which
// the pass's own stores might alias as far as the compiler knows, keep more values live and
// spill.  The explicit argument starts the segment.
struct XGroupArgs {
  XGroupDesc d[kXGroupMax];
  int nd;
};
using xgroup_args_ptr = const __attribute__((address_space(4))) XGroupArgs*;

template <typename T, int D>
__global__ void __launch_bounds__(xpass_max_threads<D>()) __attribute__((amdgpu_waves_per_eu(4)))
xpass_group_kernel(XGroupArgs args) {
  const xgroup_args_ptr ka = (xgroup_args_ptr)__builtin_amdgcn_kernarg_segment_ptr();
  int k = 0;
  for (int i = 1; i < ka->nd; ++i)
    if ((int)blockIdx.x >= ka->d[i].block0) k = i;
  const XGroupDesc& x = *(const XGroupDesc*)&ka->d[k];
  const int local = (int)blockIdx.x - x.block0;
  const int bx = local % x.gx, by = local / x.gx;
  switch (x.g.n[0]) {
    case 32: xpass_group_tile<T, 32, D>(x, bx, by); break;
    case 64: xpass_group_tile<T, 64, D>(x, bx, by); break;
    case 128: xpass_group_tile<T, 128, D>(x, bx, by); break;
    case 256: xpass_group_tile<T, 256, D>(x, bx, by); break;
    default: break;
  }
}

// descs: host array of nd <= kXGroupMax descriptors, block0 counted from 0
template <typename T, int D>
hipError_t xgroup_launch_d(hipStream_t st, const XGroupDesc* descs, int nd, int blocks,
                           int threads, size_t lds) {
  if (nd < 1 || nd > kXGroupMax) return hipErrorInvalidValue;
  XGroupArgs a{};
  for (int i = 0; i < nd; ++i) a.d[i] = descs[i];
  a.nd = nd;
  if (lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void*)xpass_group_kernel<T, D>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((xpass_group_kernel<T, D>), dim3(blocks), dim3(threads), lds, st, a);
  return hipGetLastError();
}
template hipError_t xgroup_launch_d<float, BEC_PART>(hipStream_t, const XGroupDesc*, int, int,
                                                     int, size_t);
template hipError_t xgroup_launch_d<double, BEC_PART>(hipStream_t, const XGroupDesc*, int, int,
                                                      int, size_t);

#else
extern template hipError_t xgroup_launch_d<float, 1>(hipStream_t, const XGroupDesc*, int, int, int, size_t);
extern template hipError_t xgroup_launch_d<double, 1>(hipStream_t, const XGroupDesc*, int, int, int, size_t);
extern template hipError_t xgroup_launch_d<float, 3>(hipStream_t, const XGroupDesc*, int, int, int, size_t);
extern template hipError_t xgroup_launch_d<double, 3>(hipStream_t, const XGroupDesc*, int, int, int, size_t);
extern template hipError_t xgroup_launch_d<float, 5>(hipStream_t, const XGroupDesc*, int, int, int, size_t);
extern template hipError_t xgroup_launch_d<double, 5>(hipStream_t, const XGroupDesc*, int, int, int, size_t);

bool xgroup_ok(const Geom& g) {
  const int nx = g.n[0];
  if (g.D != 1 && g.D != 3 && g.D != 5) return false;
  if (nx != 32 && nx != 64 && nx != 128 && nx != 256) return false;
  return !xpass_generic(g);
}

template <typename T>
hipError_t launch_xpass_group(hipStream_t st, int D, const XGroupDesc* descs, int nd, int blocks,
                              int threads, size_t lds) {
  switch (D) {
    case 1: return xgroup_launch_d<T, 1>(st, descs, nd, blocks, threads, lds);
    case 3: return xgroup_launch_d<T, 3>(st, descs, nd, blocks, threads, lds);
    case 5: return xgroup_launch_d<T, 5>(st, descs, nd, blocks, threads, lds);
    default: return hipErrorInvalidValue;
  }
}
template hipError_t launch_xpass_group<float>(hipStream_t, int, const XGroupDesc*, int, int, int,
                                              size_t);
template hipError_t launch_xpass_group<double>(hipStream_t, int, const XGroupDesc*, int, int, int,
                                               size_t);

// ---------------------------------------------------------------------------
// y / z pass
// ---------------------------------------------------------------------------
template <typename T, int N, int MODE>
__device__ __forceinline__ void strided_group_tile(const SGroupDesc& s, int bx_, int by_,
                                                   int bz_) {
  constexpr int TB = 2048 / N;
  constexpr int FIELDS = 1;
  const Geom& g = s.g;
  const DevTables& tb = s.tb;
  cplx<T>* __restrict__ data = static_cast<cplx<T>*>(s.data);
  const int axis = s.axis;
  const double kin_frac = s.kin_frac, dt_override = s.dt_override;
  const int itp = s.itp;
  const int ddi_data = 0;
  const int gx_ = s.gx;
#define BEC_BX bx_
#define BEC_BY by_
#define BEC_BZ bz_
#define BEC_GX gx_
#include "bec_strided_body.h"
#undef BEC_BX
#undef BEC_BY
#undef BEC_BZ
#undef BEC_GX
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kSGroupThreads)
strided_group_kernel(const SGroupDesc* __restrict__ descs, int nd) {
  const SGroupDesc& s = group_desc(descs, nd);
  const int local = (int)blockIdx.x - s.block0;
  const int bx = local % s.gx, r = local / s.gx;
  const int by = r % s.gy, bz = r / s.gy;
  switch (s.g.n[s.axis]) {
    case 32: strided_group_tile<T, 32, MODE>(s, bx, by, bz); break;
    case 64: strided_group_tile<T, 64, MODE>(s, bx, by, bz); break;
    case 128: strided_group_tile<T, 128, MODE>(s, bx, by, bz); break;
    case 256: strided_group_tile<T, 256, MODE>(s, bx, by, bz); break;
    default: break;
  }
}

bool sgroup_ok(const Geom& g, int axis, int mode, int ddi_data) {
  if (ddi_data || (mode != SM_FWD && mode != SM_INV && mode != SM_KIN)) return false;
  const int N = g.n[axis];
  return N == 32 || N == 64 || N == 128 || N == 256;
}

template <typename T>
hipError_t launch_strided_group(hipStream_t st, int mode, const SGroupDesc* descs, int nd,
                                int blocks) {
  // LDS of the largest line: 2048 elements, phase tables of N + TB + 1 entries (N + TB is at
  // most 288), reduction scratch and N twiddles
  const size_t lds = sizeof(cplx<T>) * 2048 + sizeof(double) * (2 * (256 + 64 + 1) + 16) +
                     sizeof(cplx<double>) * 256;
  switch (mode) {
    case SM_FWD:
      hipLaunchKernelGGL((strided_group_kernel<T, SM_FWD>), dim3(blocks), dim3(kSGroupThreads),
                         lds, st, descs, nd);
      break;
    case SM_INV:
      hipLaunchKernelGGL((strided_group_kernel<T, SM_INV>), dim3(blocks), dim3(kSGroupThreads),
                         lds, st, descs, nd);
      break;
    case SM_KIN:
      hipLaunchKernelGGL((strided_group_kernel<T, SM_KIN>), dim3(blocks), dim3(kSGroupThreads),
                         lds, st, descs, nd);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template hipError_t launch_strided_group<float>(hipStream_t, int, const SGroupDesc*, int, int);
template hipError_t launch_strided_group<double>(hipStream_t, int, const SGroupDesc*, int, int);
#endif

}  // namespace bec
