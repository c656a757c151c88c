// Split-step kernels for gfx950.
//
// The step is organised in HBM passes over psi (layout [instance][component][z][y][x],
// x fastest -- the reference's column-major (x, y, z, component) array):
//   x pass  : whole x rows of ALL components in LDS -> inverse x FFT, every per-site
//             operator of the potential half step (diagonal, spin mixing, DDI rotation,
//             ITP scale), spin density + R2C for the dipolar convolution, norm/energy
//             partials, forward x FFT.  One read and one write of psi.
//   y/z pass: bec_strided.hip.
// Reductions are per-block partials (fixed order) summed by a separate reduce launch,
// so results do not depend on what else runs on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "bec_device.h"
#include "bec_kernels.h"
#include "bec_passes.h"

namespace bec {

// ---------------------------------------------------------------------------
// x pass
// ---------------------------------------------------------------------------
// Waves per SIMD the x pass is compiled for.  A thread holds the E values of its FFT block
// in registers (E complex values: 4E VGPRs in FP64), so lines with E = 20 or 28 (5 2^k and
// 7 2^k) and lines with E = 12 (3 2^k) get a larger register budget; at 4 waves (128 VGPRs)
// they spilled hundreds of registers to scratch.
template <typename T, int NX, int D>
constexpr int xpass_waves() {
  constexpr int e = line_e(NX);
  if (e >= 20) return 2;
  if (e == 12) return 3;
  return D == 13 && sizeof(T) == 8 ? 3 : 4;
}

template <typename T, int NX, int D>
__global__ void __launch_bounds__(xpass_max_threads<D>())
    __attribute__((amdgpu_waves_per_eu(xpass_waves<T, NX, D>())))
xpass_kernel(Geom g, DevTables tb, XProgram prog, const cplx<T>* __restrict__ src,
             cplx<T>* __restrict__ dst, cplx<T>* __restrict__ fhat, T* __restrict__ phi, int ROWS) {
#define BEC_BX blockIdx.x
#define BEC_BY blockIdx.y
#define BEC_REPEAT 0
#include "bec_xpass_body.h"
#undef BEC_REPEAT
#undef BEC_BX
#undef BEC_BY
}

// The same pass with prog.nrep split steps in one launch (1-D lines of F = 0 and 1, whose
// split step is this one pass; the schedule compiler decides, bec_split_step).  A kernel of
// its own, so the single-step kernel's code is not touched by the repetition loop.  One
// block per instance of max(D NX / E, NX) threads: one thread per site for the per-site
// operators, whose dependent chains (exp, sinh, cosh, sqrt) set the step's latency on a
// single block.  Occupancy does not matter, so the kernel may use the whole register file.
template <typename T, int NX, int D>
__global__ void __launch_bounds__(xpass_max_threads<D>()) __attribute__((amdgpu_waves_per_eu(1)))
xpass_rep_kernel(Geom g, DevTables tb, XProgram prog, const cplx<T>* __restrict__ src,
                 cplx<T>* __restrict__ dst, cplx<T>* __restrict__ fhat, T* __restrict__ phi,
                 int ROWS) {
#define BEC_BX blockIdx.x
#define BEC_BY blockIdx.y
#define BEC_REPEAT 1
#include "bec_xpass_body.h"
#undef BEC_REPEAT
#undef BEC_BX
#undef BEC_BY
}

// ---------------------------------------------------------------------------
// pointwise energy terms (energy.jl:6-121, observables.jl:1-33): per-block partials of
// norm, trap, Zeeman, n^2, |F|^2, n^{5/2}, Phi.F, Mz and the tensor pair-amplitude sum.
// ---------------------------------------------------------------------------
template <typename T, int D>
__global__ void __launch_bounds__(256)
energy_kernel(Geom g, DevTables tb, const cplx<T>* __restrict__ psi, const T* __restrict__ phi,
              int with_tensor) {
  __shared__ double rsh[16];
  constexpr int F = (D - 1) / 2;
  const int inst = tb.active[blockIdx.y];
  const InstParams& P = tb.params[inst];
  double e_norm = 0, e_trap = 0, e_zee = 0, e_c0 = 0, e_c1 = 0, e_lhy = 0, e_ddi = 0, e_mz = 0;
  double e_ten = 0, e_ram = 0;
  const cplx<T>* p = psi + (size_t)inst * g.inst_stride;
  for (long long site = blockIdx.x * (long long)blockDim.x + threadIdx.x; site < g.nsites;
       site += (long long)gridDim.x * blockDim.x) {
    cplx<T> sp[D];
#pragma unroll
    for (int c = 0; c < D; ++c) sp[c] = p[(size_t)c * g.nsites + site];
    double n = 0, zz = 0, mz = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double a2 = (double)abs2(sp[c]);
      n += a2;
      zz += P.zee[c] * a2;
      mz += (double)(F - c) * a2;
    }
    T fx, fy, fz;
    spin_vec<T, D>(sp, fx, fy, fz);
    e_norm += n;
    e_trap += trap_site(g, tb, inst, site) * n;
    e_zee += zz;
    e_mz += mz;
    e_c0 += n * n;
    e_c1 += (double)fx * fx + (double)fy * fy + (double)fz * fz;
    e_lhy += n * n * sqrt(n);
    if (P.has_raman) {  // _raman_energy (energy.jl:123-135): psi^H H_R psi = phi . <F>
      double phx, phy, phz;
      raman_phi(g, tb, P, site, phx, phy, phz);
      e_ram += phx * (double)fx + phy * (double)fy + phz * (double)fz;
    }
    if (with_tensor) {
      if constexpr (D >= 3) e_ten += tensor_energy_site<T, D>(sp, P);
    } else if (P.c2 != 0.0) {  // energy.jl:43-46,106-109: 1/2 c2 |A00|^2, tensor slot
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const double sg = (c % 2 == 0) ? 1.0 : -1.0;
        const double xr = sp[c].x, xi = sp[c].y, yr = sp[D - 1 - c].x, yi = sp[D - 1 - c].y;
        ar += sg * (xr * yr - xi * yi);
        ai += sg * (xr * yi + xi * yr);
      }
      e_ten += 0.5 * P.c2 * (ar * ar + ai * ai) / (double)D;
    }
    if (phi) {
      const T* ph = phi + (size_t)inst * 3 * g.nsites + site;
      e_ddi += (double)ph[0] * fx + (double)ph[g.nsites] * fy + (double)ph[2 * g.nsites] * fz;
    }
  }
  const int blk = blockIdx.x, tid = threadIdx.x;
  double v;
  v = block_sum(e_norm, rsh); if (tid == 0) *partial_ptr(tb, RS_NORM, inst, blk) = v;
  v = block_sum(e_trap, rsh); if (tid == 0) *partial_ptr(tb, RS_TRAP, inst, blk) = v;
  v = block_sum(e_zee, rsh); if (tid == 0) *partial_ptr(tb, RS_ZEE, inst, blk) = v;
  v = block_sum(e_c0, rsh); if (tid == 0) *partial_ptr(tb, RS_C0, inst, blk) = v;
  v = block_sum(e_c1, rsh); if (tid == 0) *partial_ptr(tb, RS_C1, inst, blk) = v;
  v = block_sum(e_lhy, rsh); if (tid == 0) *partial_ptr(tb, RS_LHY, inst, blk) = v;
  v = block_sum(e_ddi, rsh); if (tid == 0) *partial_ptr(tb, RS_DDI, inst, blk) = v;
  v = block_sum(e_mz, rsh); if (tid == 0) *partial_ptr(tb, RS_MZ, inst, blk) = v;
  v = block_sum(e_ten, rsh); if (tid == 0) *partial_ptr(tb, RS_TENSOR, inst, blk) = v;
  v = block_sum(e_ram, rsh); if (tid == 0) *partial_ptr(tb, RS_RAMAN, inst, blk) = v;
}

// ---------------------------------------------------------------------------
// lane-per-site pass (13-component operators: tensor expm-action, Euler rotations)
// ---------------------------------------------------------------------------
template <typename T, int D>
__device__ __forceinline__ void site_ops(const Geom& g, const DevTables& tb, const SProgram& prog,
                                         cplx<T>* __restrict__ psi, const T* __restrict__ phi,
                                         const Op* s_ops, const cplx<double>* zf, int inst,
                                         long long site, double dt, bool itp, int& nmv, int& nrot) {
  const InstParams& P = tb.params[inst];
  cplx<T>* p = psi + (size_t)inst * g.inst_stride + site;
  cplx<T> sp[D];
#pragma unroll
  for (int c = 0; c < D; ++c) sp[c] = p[(size_t)c * g.nsites];
  for (int i = 0; i < prog.n; ++i) {
    const Op op = s_ops[i];
    const double dtf = op.frac * dt;
    if (op.kind == OP_DIAG) {
      const double Vt = trap_site(g, tb, inst, site);
      op_diag<T, D>(sp, Vt, P, dtf, itp, &zf[i * D]);
    } else if (op.kind == OP_SPINMIX) {
      nrot += op_spin_mixing<T, D>(sp, P, dtf, itp);
    } else if (op.kind == OP_DDI_ROT) {
      const T* ph = phi + (size_t)inst * 3 * g.nsites + site;
      nrot += op_euler<T, D>(sp, (double)ph[0], (double)ph[g.nsites], (double)ph[2 * g.nsites], dtf,
                             itp);
    } else if (op.kind == OP_NEMATIC) {
      op_nematic<T, D>(sp, P.c2, dtf, itp);
    } else if (op.kind == OP_LOSS) {
      op_loss<T, D>(sp, P, dtf);
    } else if (op.kind == OP_RAMAN) {
      if (P.has_raman) {
        double phx, phy, phz;
        raman_phi(g, tb, P, site, phx, phy, phz);
        nrot += op_euler<T, D>(sp, phx, phy, phz, dtf, itp);
      }
    } else if (op.kind == OP_TENSOR) {
      if constexpr (D >= 3)
        nmv += op_tensor<T, D>(sp, P, dtf, itp, tb.kfold + (size_t)inst * tens::kKfoldMax);
    } else if (op.kind == OP_SCALE) {
      const double sc = 1.0 / sqrt(tb.red[(size_t)inst * RS_COUNT + RS_NORM] * g.dV);
#pragma unroll
      for (int c = 0; c < D; ++c) sp[c] = scale64(sp[c], sc);
    }
  }
#pragma unroll
  for (int c = 0; c < D; ++c) p[(size_t)c * g.nsites] = sp[c];
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
site_kernel(Geom g, DevTables tb, SProgram prog, cplx<T>* __restrict__ psi,
            const T* __restrict__ phi) {
  __shared__ cplx<double> zf[kMaxOps * D];
  __shared__ Op s_ops[kMaxOps];
  const int tid = threadIdx.x;
  const int inst = tb.active[blockIdx.y];
  const InstParams& P = tb.params[inst];
  const double dt = prog.dt_override > 0.0 ? prog.dt_override : P.dt;
  const bool itp = prog.itp != 0;
#pragma unroll
  for (int i = 0; i < kMaxOps; ++i)
    if (tid == (i < (int)blockDim.x ? i : 0)) s_ops[i] = prog.ops[i];
  __syncthreads();
  for (int idx = tid; idx < kMaxOps * D; idx += blockDim.x) {
    const int i = idx / D, c = idx % D;
    const Op op = s_ops[i];
    if (i < prog.n && op.kind == OP_DIAG) {
      const double dtf = op.frac * dt;
      if (itp) zf[idx] = mk(exp(-(P.zee[c] - P.zee_shift) * dtf), 0.0);
      else zf[idx] = cis_neg<double>(P.zee[c] * dtf);
    }
  }
  __syncthreads();
  const long long site = (long long)blockIdx.x * blockDim.x + tid;
  int nmv = 0, nrot = 0;
  if (site < g.nsites) site_ops<T, D>(g, tb, prog, psi, phi, s_ops, zf, inst, site, dt, itp, nmv, nrot);
  if (tb.work) {  // profiling: executed tensor matvecs and spin rotations
    // wave sums (no LDS, no barrier), then one atomic per wave into a spread slot
    // matvecs, rotations, first matvecs in the symmetric form (op_tensor's tag bits)
    int a = nmv & ((1 << kSymShift) - 1), b = nrot, c = nmv >> kSymShift;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o);
      b += __shfl_xor(b, o);
      c += __shfl_xor(c, o);
    }
    if ((tid & 63) == 0) {
      const unsigned slot = (blockIdx.x * (blockDim.x >> 6) + (tid >> 6)) % kWorkSlots;
      unsigned long long* w = tb.work + (size_t)slot * kWorkStride;
      if (a) atomicAdd(&w[0], (unsigned long long)a);
      if (b) atomicAdd(&w[1], (unsigned long long)b);
      if (c) atomicAdd(&w[2], (unsigned long long)c);
    }
  }
}


#if BEC_PART == 0  // shape-independent kernels: in the shared part only
// ---------------------------------------------------------------------------
// reductions and small passes
// ---------------------------------------------------------------------------
__global__ void reduce_kernel(DevTables tb, double* red, int nblocks, uint32_t slot_mask,
                              uint32_t max_mask) {
  __shared__ double sh[16];
  // blockIdx.x enumerates the set bits of slot_mask
  int slot = -1, cnt = 0;
  for (int s = 0; s < RS_COUNT; ++s)
    if (slot_mask & (1u << s)) {
      if (cnt == (int)blockIdx.x) { slot = s; break; }
      ++cnt;
    }
  if (slot < 0) return;
  const int inst = tb.active[blockIdx.y];
  const double* p = tb.partials + ((size_t)slot * tb.maxB + inst) * tb.max_blocks;
  const bool is_max = (max_mask >> slot) & 1u;
  double acc = 0.0, comp = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
    const double v = p[i];
    if (is_max) {
      acc = fmax(acc, v);
    } else {  // Neumaier compensated summation
      const double t = acc + v;
      comp += fabs(acc) >= fabs(v) ? (acc - t) + v : (v - t) + acc;
      acc = t;
    }
  }
  double r;
  if (is_max) {
    r = block_max(acc, sh);
  } else {
    const double a = block_sum(acc, sh);
    __syncthreads();
    const double c = block_sum(comp, sh);
    r = a + c;
  }
  if (threadIdx.x == 0) red[(size_t)inst * RS_COUNT + slot] = r;
}

template <typename T>
__global__ void scale_norm_kernel(Geom g, DevTables tb, cplx<T>* psi) {
  const int inst = tb.active[blockIdx.y];
  const double n2 = tb.red[(size_t)inst * RS_COUNT + RS_NORM] * g.dV;
  const double s = 1.0 / sqrt(n2);
  cplx<T>* p = psi + (size_t)inst * g.inst_stride;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.inst_stride;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = scale64(p[i], s);
}

// Grouped forms of the reduction and of the renormalisation (bec_group.hip): block b of the
// launch runs block (b - block0) of its descriptor's own launch shape.
template <typename Desc>
__device__ __forceinline__ const Desc& group_desc0(const Desc* d, int nd) {
  int k = 0;
  for (int i = 1; i < nd; ++i)
    if ((int)blockIdx.x >= d[i].block0) k = i;
  return d[k];
}
__global__ void reduce_group_kernel(const RGroupDesc* __restrict__ descs, int nd) {
  __shared__ double sh[16];
  const RGroupDesc& r = group_desc0(descs, nd);
  const int local = (int)blockIdx.x - r.block0;
  const int bx = local % r.gx, by = local / r.gx;
  int slot = -1, cnt = 0;
  for (int s = 0; s < RS_COUNT; ++s)
    if (r.slot_mask & (1u << s)) {
      if (cnt == bx) { slot = s; break; }
      ++cnt;
    }
  if (slot < 0) return;
  const DevTables& tb = r.tb;
  const int inst = tb.active[by];
  const double* p = tb.partials + ((size_t)slot * tb.maxB + inst) * tb.max_blocks;
  const bool is_max = (r.max_mask >> slot) & 1u;
  double acc = 0.0, comp = 0.0;
  for (int i = threadIdx.x; i < r.nblocks; i += blockDim.x) {
    const double v = p[i];
    if (is_max) {
      acc = fmax(acc, v);
    } else {  // Neumaier compensated summation, as reduce_kernel
      const double t = acc + v;
      comp += fabs(acc) >= fabs(v) ? (acc - t) + v : (v - t) + acc;
      acc = t;
    }
  }
  double res;
  if (is_max) {
    res = block_max(acc, sh);
  } else {
    const double a = block_sum(acc, sh);
    __syncthreads();
    const double c = block_sum(comp, sh);
    res = a + c;
  }
  if (threadIdx.x == 0) r.red[(size_t)inst * RS_COUNT + slot] = res;
}

template <typename T>
__global__ void scale_group_kernel(const CGroupDesc* __restrict__ descs, int nd) {
  const CGroupDesc& c = group_desc0(descs, nd);
  const int local = (int)blockIdx.x - c.block0;
  const int bx = local % c.gx, by = local / c.gx;
  const int inst = c.tb.active[by];
  const double n2 = c.tb.red[(size_t)inst * RS_COUNT + RS_NORM] * c.g.dV;
  const double s = 1.0 / sqrt(n2);
  cplx<T>* p = static_cast<cplx<T>*>(c.psi) + (size_t)inst * c.g.inst_stride;
  for (long long i = bx * (long long)blockDim.x + threadIdx.x; i < c.g.inst_stride;
       i += (long long)c.gx * blockDim.x)
    p[i] = scale64(p[i], s);
}

// Per-block, per-component partial norms: part[(inst*D + c)*nblk + blk] = sum |psi_c|^2.
template <typename T>
__global__ void comp_norm_kernel(Geom g, DevTables tb, const cplx<T>* __restrict__ psi,
                                 double* __restrict__ part) {
  __shared__ double sh[16];
  const int inst = tb.active[blockIdx.y];
  const cplx<T>* p = psi + (size_t)inst * g.inst_stride;
  for (int c = 0; c < g.D; ++c) {
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.nsites;
         i += (long long)gridDim.x * blockDim.x)
      acc += (double)abs2(p[(size_t)c * g.nsites + i]);
    const double v = block_sum(acc, sh);
    if (threadIdx.x == 0) part[((size_t)inst * g.D + c) * gridDim.x + blockIdx.x] = v;
  }
}

// ground_state.jl:279-320 for one instance per block (thread 0): component norms N_c
// (fixed-order compensated sums of the partials, times dV), then the factors of
// normalise -> Newton on lambda for Mz(lambda) = target (<= 20 iterations, lambda clamped
// to [-10, 10], weights exp(lambda m)) -> normalise, composed into one factor per component.
__global__ void constrained_factor_kernel(Geom g, DevTables tb, const double* __restrict__ part,
                                          int nblk, const double* __restrict__ target,
                                          double* __restrict__ fac) {
  if (threadIdx.x != 0) return;
  const int inst = tb.active[blockIdx.x];
  const int D = g.D, F = (D - 1) / 2;
  double nc[kMaxD];
  double tot = 0.0;
  for (int c = 0; c < D; ++c) {
    double s = 0.0, comp = 0.0;
    const double* q = part + ((size_t)inst * D + c) * nblk;
    for (int b = 0; b < nblk; ++b) {  // Neumaier
      const double t = s + q[b];
      comp += fabs(s) >= fabs(q[b]) ? (s - t) + q[b] : (q[b] - t) + s;
      s = t;
    }
    nc[c] = (s + comp) * g.dV;
    tot += nc[c];
  }
  const double s0 = tot > 0.0 ? 1.0 / sqrt(tot) : 1.0;  // first _normalize_psi!
  for (int c = 0; c < D; ++c) nc[c] *= s0 * s0;
  const double tz = target[inst];
  double lambda = 0.0;
  for (int it = 0; it < 20; ++it) {
    double w2n[kMaxD], total = 0.0;
    for (int c = 0; c < D; ++c) {
      const double w = exp(lambda * (double)(F - c));
      w2n[c] = nc[c] * w * w;
      total += w2n[c];
    }
    if (total < 1e-30) break;
    double mz = 0.0;
    for (int c = 0; c < D; ++c) mz += (double)(F - c) * w2n[c];
    mz /= total;
    if (fabs(mz - tz) < 1e-12) break;
    double dmz = 0.0;
    for (int c = 0; c < D; ++c) {
      const double m = (double)(F - c);
      dmz += 2.0 * m * (m - mz) * w2n[c] / total;
    }
    if (fabs(dmz) < 1e-30) break;
    lambda -= (mz - tz) / dmz;
    lambda = lambda < -10.0 ? -10.0 : (lambda > 10.0 ? 10.0 : lambda);
  }
  double tw = 0.0;
  double w[kMaxD];
  for (int c = 0; c < D; ++c) {
    w[c] = exp(lambda * (double)(F - c));
    tw += nc[c] * w[c] * w[c];
  }
  const double s1 = tw > 0.0 ? 1.0 / sqrt(tw) : 1.0;  // final _normalize_psi!
  for (int c = 0; c < D; ++c) fac[(size_t)inst * kMaxD + c] = s0 * w[c] * s1;
}

template <typename T>
__global__ void comp_scale_kernel(Geom g, DevTables tb, cplx<T>* psi,
                                  const double* __restrict__ fac) {
  const int inst = tb.active[blockIdx.y];
  cplx<T>* p = psi + (size_t)inst * g.inst_stride;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.inst_stride;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = scale64(p[i], fac[(size_t)inst * kMaxD + i / g.nsites]);
}

template <typename T>
__global__ void diff_kernel(Geom g, DevTables tb, const cplx<T>* __restrict__ a,
                            const cplx<T>* __restrict__ b, int want_max) {
  __shared__ double sh[16];
  const int inst = tb.active[blockIdx.y];
  const cplx<T>* pa = a + (size_t)inst * g.inst_stride;
  const cplx<T>* pb = b + (size_t)inst * g.inst_stride;
  double s0 = 0.0, s1 = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.inst_stride;
       i += (long long)gridDim.x * blockDim.x) {
    const cplx<T> va = pa[i], vb = pb[i];
    const double dr = (double)va.x - (double)vb.x, di = (double)va.y - (double)vb.y;
    const double d2 = dr * dr + di * di;
    if (want_max) {
      s0 = fmax(s0, sqrt(d2));
      s1 = fmax(s1, sqrt((double)abs2(va)));
    } else {
      s0 += d2;
      s1 += (double)abs2(vb);
    }
  }
  double r0, r1;
  if (want_max) {
    r0 = block_max(s0, sh);
    __syncthreads();
    r1 = block_max(s1, sh);
  } else {
    r0 = block_sum(s0, sh);
    __syncthreads();
    r1 = block_sum(s1, sh);
  }
  if (threadIdx.x == 0) {
    *partial_ptr(tb, RS_NORM, inst, blockIdx.x) = r0;
    *partial_ptr(tb, RS_KIN, inst, blockIdx.x) = r1;
  }
}

#endif  // BEC_PART == 0

// ---------------------------------------------------------------------------
// launchers
//
// This file is compiled once per part: BEC_PART = 0 builds the shape-independent code (the
// dispatchers, reductions, small passes); BEC_PART = D builds the x-pass, energy and
// lane-per-site kernels of D = 2F+1 components only, so the parts compile in parallel.
// ---------------------------------------------------------------------------
#ifndef BEC_PART
#error "compile with -DBEC_PART=0 (shared code) or -DBEC_PART=D (kernels of D components)"
#endif

static size_t xpass_lds(int D, int rows, int NX, size_t csz, bool has_aux, int AL) {
  // padded tiles: 9 slots per 8 elements; AL: length of the DDI auxiliary lines
  // (the tile part rounded up to an even element count, as the kernel lays it out)
  return csz * (((size_t)D * rows * NX / 8 * 9 + 1) / 2 * 2 +
                (has_aux ? 2 * (size_t)rows * AL / 8 * 9 : 0)) +
         sizeof(cplx<double>) * 2 * kMaxOps * D + 16 * sizeof(double) +
         sizeof(cplx<double>) * NX;  // twiddles
}

template <typename T, int NX, int D>
static hipError_t xpass_n(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                          const void* src, void* dst, void* fhat, int n_active, void* phi) {
  const int rows = xpass_rows_g(g);
  const bool has_aux = (prog.stages & (XS_LOAD_PHI | XS_WRITE_FHAT)) != 0;
  const size_t lds = xpass_lds(D, rows, NX, sizeof(cplx<T>), has_aux, g.spad ? 2 * NX : NX);
  const int threads = D * rows * (NX / xpass_e<NX, D>());
  if (threads > xpass_max_threads<D>()) return hipErrorInvalidConfiguration;
  dim3 grid(xpass_grid(g), n_active);
  if (prog.nrep > 1) {
    if constexpr (D <= 3) {
      if (lds > 48 * 1024)
        (void)hipFuncSetAttribute((const void*)xpass_rep_kernel<T, NX, D>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      // one thread per site for the per-site operators (the tile's FFT threads are the first
      // D rows NX / E of them)
      const int rthreads = std::max(threads, rows * NX);
      if (rthreads > xpass_max_threads<D>()) return hipErrorInvalidConfiguration;
      hipLaunchKernelGGL((xpass_rep_kernel<T, NX, D>), grid, dim3(rthreads), lds, st, g, tb, prog,
                         (const cplx<T>*)src, (cplx<T>*)dst, (cplx<T>*)fhat, (T*)phi, rows);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (lds > 48 * 1024)  // + static LDS (op lists)
    (void)hipFuncSetAttribute((const void*)xpass_kernel<T, NX, D>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((xpass_kernel<T, NX, D>), grid, dim3(threads), lds, st, g, tb, prog,
                     (const cplx<T>*)src, (cplx<T>*)dst, (cplx<T>*)fhat, (T*)phi, rows);
  return hipGetLastError();
}

template <typename T, int D>
hipError_t xpass_launch_d(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                          const void* src, void* dst, void* fhat, int n_active, void* phi) {
  switch (g.n[0]) {
    case 8: return xpass_n<T, 8, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 24: return xpass_n<T, 24, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 40: return xpass_n<T, 40, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 48: return xpass_n<T, 48, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 80: return xpass_n<T, 80, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 96: return xpass_n<T, 96, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 160: return xpass_n<T, 160, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 192: return xpass_n<T, 192, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 56: return xpass_n<T, 56, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 112: return xpass_n<T, 112, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 224: return xpass_n<T, 224, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 16: return xpass_n<T, 16, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 32: return xpass_n<T, 32, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 64: return xpass_n<T, 64, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 128: return xpass_n<T, 128, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 256: return xpass_n<T, 256, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    case 512: return xpass_n<T, 512, D>(st, g, tb, prog, src, dst, fhat, n_active, phi);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int D>
hipError_t energy_launch_d(hipStream_t st, const Geom& g, const DevTables& tb, const void* psi,
                           const void* phi, int with_tensor, dim3 grid) {
  hipLaunchKernelGGL((energy_kernel<T, D>), grid, dim3(256), 0, st, g, tb, (const cplx<T>*)psi,
                     (const T*)phi, with_tensor);
  return hipGetLastError();
}

template <typename T, int D>
hipError_t sites_launch_d(hipStream_t st, const Geom& g, const DevTables& tb, const SProgram& prog,
                          void* psi, const void* phi, dim3 grid) {
  hipLaunchKernelGGL((site_kernel<T, D>), grid, dim3(256), 0, st, g, tb, prog, (cplx<T>*)psi,
                     (const T*)phi);
  return hipGetLastError();
}

#define BEC_D_INST(PRE, T, D)                                                                 \
  PRE template hipError_t xpass_launch_d<T, D>(hipStream_t, const Geom&, const DevTables&,     \
                                               const XProgram&, const void*, void*, void*, int,\
                                               void*);                                         \
  PRE template hipError_t energy_launch_d<T, D>(hipStream_t, const Geom&, const DevTables&,    \
                                                const void*, const void*, int, dim3);          \
  PRE template hipError_t sites_launch_d<T, D>(hipStream_t, const Geom&, const DevTables&,     \
                                               const SProgram&, void*, const void*, dim3);

#if BEC_PART == 0
// the per-D launchers live in the other parts
#define BEC_D_EXTERN(D) BEC_D_INST(extern, float, D) BEC_D_INST(extern, double, D)
BEC_D_EXTERN(1)
BEC_D_EXTERN(3)
BEC_D_EXTERN(5)
BEC_D_EXTERN(7)
BEC_D_EXTERN(9)
BEC_D_EXTERN(11)
BEC_D_EXTERN(13)

static int xpass_tile_elems() {
  static int v = [] {
    const char* e = getenv("BEC_XPASS_TILE");
    const int t = e ? atoi(e) : 0;
    return t >= 256 ? t : 1536;
  }();
  return v;
}

int xpass_rows(int D, int NX, int nrows) {
  // the element budget is for 8 values per thread; lines with E = 12, 20 or 28 take up to
  // twice it so that their blocks keep about the same thread count (at most twice: an FP64
  // spin-1 tile with zero-padded DDI lines then still fits in LDS, about 130 KB)
  const int e = line_e(NX);
  int target = xpass_tile_elems() * std::min(e, 16) / 8 / (D * NX);
  const int maxthr = 512;
  while (target > 1 && D * target * (NX / e) > maxthr) --target;
  if (target < 1) target = 1;
  int r = 1;
  while (r * 2 <= target) r *= 2;
  while (r > 1 && (nrows % r) != 0) r /= 2;
  if (r > nrows) r = nrows;
  // LDS bound for the largest program (FP64, zero-padded DDI lines)
  while (r > 1 && xpass_lds(D, r, NX, 16, true, 2 * NX) > 160 * 1024) r /= 2;
  return r;
}


int xpass_rows_g(const Geom& g) {
  if (g.xrows > 0) return g.xrows;
  return xpass_generic(g) ? xgen_rows(g) : xpass_rows(g.D, g.n[0], g.nrows);
}

bool xpass_rows_ok(const Geom& g, int rows) {
  // the LDS bound is taken for FP64 with the DDI lines, so one choice fits every program
  // (the runtime-length x pass keeps its own partition)
  const int AL = g.spad ? 2 * g.n[0] : g.n[0];
  return !xpass_generic(g) && rows >= 1 && rows <= g.nrows && g.nrows % rows == 0 &&
         g.D * rows * (g.n[0] / line_e(g.n[0])) <= 512 &&
         xpass_lds(g.D, rows, g.n[0], sizeof(cplx<double>), true, AL) <= 160 * 1024;
}

int xpass_grid(const Geom& g) { return g.nrows / xpass_rows_g(g); }

// Component counts with built kernels: every spin F = 0 .. 6.
#define BEC_D_SWITCH(CALL)       \
  switch (g.D) {                 \
    case 1: return CALL(1);      \
    case 3: return CALL(3);      \
    case 5: return CALL(5);      \
    case 7: return CALL(7);      \
    case 9: return CALL(9);      \
    case 11: return CALL(11);    \
    case 13: return CALL(13);    \
    default: return hipErrorInvalidValue; \
  }

template <typename T>
hipError_t launch_xpass(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                        const void* src, void* dst, void* fhat, int n_active, void* phi) {
  if (xpass_generic(g)) return launch_xpass_gen<T>(st, g, tb, prog, src, dst, fhat, n_active, phi);
#define BEC_CALL(D) xpass_launch_d<T, D>(st, g, tb, prog, src, dst, fhat, n_active, phi)
  BEC_D_SWITCH(BEC_CALL)
#undef BEC_CALL
}

template <typename T>
hipError_t launch_energy(hipStream_t st, const Geom& g, const DevTables& tb, const void* psi,
                         const void* phi, int with_tensor, int n_active, int* nblocks_out) {
  long long nb = (g.nsites + 255) / 256;
  if (nb > 1024) nb = 1024;
  *nblocks_out = (int)nb;
  const dim3 grid((unsigned)nb, n_active);
#define BEC_CALL(D) energy_launch_d<T, D>(st, g, tb, psi, phi, with_tensor, grid)
  BEC_D_SWITCH(BEC_CALL)
#undef BEC_CALL
}

template <typename T>
hipError_t launch_sites(hipStream_t st, const Geom& g, const DevTables& tb, const SProgram& prog,
                        void* psi, const void* phi, int n_active) {
  const dim3 grid((unsigned)((g.nsites + 255) / 256), n_active);
#define BEC_CALL(D) sites_launch_d<T, D>(st, g, tb, prog, psi, phi, grid)
  BEC_D_SWITCH(BEC_CALL)
#undef BEC_CALL
}

hipError_t launch_reduce(hipStream_t st, const DevTables& tb, double* red, int n_active,
                         int nblocks, uint32_t slot_mask, uint32_t max_mask) {
  int nslots = __builtin_popcount(slot_mask);
  if (nslots == 0 || n_active == 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_kernel, dim3(nslots, n_active), dim3(256), 0, st, tb, red, nblocks,
                     slot_mask, max_mask);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_norm(hipStream_t st, const Geom& g, const DevTables& tb, void* psi,
                             int n_active) {
  long long nb = (g.inst_stride + 255) / 256;
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL((scale_norm_kernel<T>), dim3((int)nb, n_active), dim3(256), 0, st, g, tb,
                     (cplx<T>*)psi);
  return hipGetLastError();
}

hipError_t launch_reduce_group(hipStream_t st, const RGroupDesc* descs, int nd, int blocks) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_group_kernel, dim3(blocks), dim3(256), 0, st, descs, nd);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_group(hipStream_t st, const CGroupDesc* descs, int nd, int blocks) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((scale_group_kernel<T>), dim3(blocks), dim3(256), 0, st, descs, nd);
  return hipGetLastError();
}
template hipError_t launch_scale_group<float>(hipStream_t, const CGroupDesc*, int, int);
template hipError_t launch_scale_group<double>(hipStream_t, const CGroupDesc*, int, int);

size_t xpass_lds_g(const Geom& g, const XProgram& prog, size_t csz) {
  const bool has_aux = (prog.stages & (XS_LOAD_PHI | XS_WRITE_FHAT)) != 0;
  return xpass_lds(g.D, xpass_rows_g(g), g.n[0], csz, has_aux, g.spad ? 2 * g.n[0] : g.n[0]);
}
int xpass_threads_g(const Geom& g) {
  return g.D * xpass_rows_g(g) * (g.n[0] / line_e(g.n[0]));
}

template <typename T>
hipError_t launch_constrained_norm(hipStream_t st, const Geom& g, const DevTables& tb, void* psi,
                                   int n_active, double* part, int nblk, const double* target,
                                   double* fac) {
  hipLaunchKernelGGL((comp_norm_kernel<T>), dim3(nblk, n_active), dim3(256), 0, st, g, tb,
                     (const cplx<T>*)psi, part);
  hipLaunchKernelGGL(constrained_factor_kernel, dim3(n_active), dim3(64), 0, st, g, tb, part,
                     nblk, target, fac);
  long long nb = (g.inst_stride + 255) / 256;
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL((comp_scale_kernel<T>), dim3((int)nb, n_active), dim3(256), 0, st, g, tb,
                     (cplx<T>*)psi, (const double*)fac);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_diff(hipStream_t st, const Geom& g, const DevTables& tb, const void* a,
                       const void* b, int n_active, int want_max, int nblocks) {
  hipLaunchKernelGGL((diff_kernel<T>), dim3(nblocks, n_active), dim3(256), 0, st, g, tb,
                     (const cplx<T>*)a, (const cplx<T>*)b, want_max);
  return hipGetLastError();
}

__global__ void widen_kernel(const float* __restrict__ src, double* __restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = (double)src[i];
}

hipError_t launch_widen(hipStream_t st, const void* src, void* dst, long long n) {
  long long nb = (n + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(widen_kernel, dim3((unsigned)nb), dim3(256), 0, st, (const float*)src,
                     (double*)dst, n);
  return hipGetLastError();
}

// Slot copy of the active instances (bec_copy_slot): 16-byte vectors, eight loads per thread
// in flight before their stores, non-temporal (the copy is read and written once; keeping it
// out of the caches leaves them to the passes around it).  Grid: (blocks, active
// instances); n16 vectors per instance.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256)
copy_slot_kernel(DevTables tb, const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                 long long n16) {
  constexpr int U = 8;
  const int inst = tb.active[blockIdx.y];
  const u32x4* s = src + (size_t)inst * n16;
  u32x4* d = dst + (size_t)inst * n16;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n16; i += U * stride) {
    u32x4 v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = __builtin_nontemporal_load(&s[i + k * stride]);
#pragma unroll
    for (int k = 0; k < U; ++k) __builtin_nontemporal_store(v[k], &d[i + k * stride]);
  }
  for (; i < n16; i += stride) d[i] = s[i];
}

hipError_t launch_copy_slot(hipStream_t st, const DevTables& tb, const void* src, void* dst,
                            long long bytes_per_inst, int n_active) {
  if (bytes_per_inst % 16 != 0) return hipErrorInvalidValue;
  const long long n16 = bytes_per_inst / 16;
  long long nb = (n16 + 2047) / 2048;
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(copy_slot_kernel, dim3((unsigned)nb, n_active), dim3(256), 0, st, tb,
                     (const u32x4*)src, (u32x4*)dst, n16);
  return hipGetLastError();
}

#define BEC_INST(T)                                                                           \
  template hipError_t launch_xpass<T>(hipStream_t, const Geom&, const DevTables&,              \
                                      const XProgram&, const void*, void*, void*, int, void*); \
  template hipError_t launch_energy<T>(hipStream_t, const Geom&, const DevTables&, const void*, \
                                       const void*, int, int, int*);                          \
  template hipError_t launch_sites<T>(hipStream_t, const Geom&, const DevTables&,              \
                                      const SProgram&, void*, const void*, int);               \
  template hipError_t launch_scale_norm<T>(hipStream_t, const Geom&, const DevTables&, void*,  \
                                           int);                                               \
  template hipError_t launch_diff<T>(hipStream_t, const Geom&, const DevTables&, const void*,  \
                                     const void*, int, int, int);                             \
  template hipError_t launch_constrained_norm<T>(hipStream_t, const Geom&, const DevTables&,   \
                                                 void*, int, double*, int, const double*,     \
                                                 double*);
BEC_INST(double)
BEC_INST(float)

#else  // BEC_PART == D: the kernels of D components
BEC_D_INST(, float, BEC_PART)
BEC_D_INST(, double, BEC_PART)
#endif

}  // namespace bec
