// Fused x pass + lane-per-site pass for spins whose per-site operators are heavy (F >= 2:
// Euler spin rotations, the general-F tensor step, the F = 6 nematic step).
//
// One block holds ROWS whole x rows of every component in LDS, ROWS*NX sites with one
// thread per site, and runs the stages the schedule asks for:
//   inverse x FFT -> C2R of the three DDI potential lines (Phi stays in LDS) -> the site
//   program on the spinor in VGPRs (diagonal, spin mixing, Raman, DDI rotation by Phi,
//   tensor, nematic, loss, ITP scale) -> spin density and its R2C into the DDI spectra -> norm
//   partials -> forward x FFT -> store.
// Without it a half potential step of the dipolar spin-6 gas is six launches (x pass,
// site pass, spin-density x pass, Phi x pass, site pass, x pass: four extra round trips
// of psi through HBM); with it, two.  The x-line work hides behind the VALU-bound tensor
// step.  Each operator is the same device function as in the x pass and the site pass,
// so a step gives the same bits whichever kernel runs it.
#include <hip/hip_runtime.h>

#include "bec_device.h"
#include "bec_kernels.h"
#include "bec_passes.h"

namespace bec {

constexpr int kXsiteMaxThreads = 256;

template <typename T, int NX, int D>
__global__ void __launch_bounds__(kXsiteMaxThreads) __attribute__((amdgpu_waves_per_eu(2)))
xsite_kernel(Geom g, DevTables tb, XProgram prog, const cplx<T>* __restrict__ src,
             cplx<T>* __restrict__ psi, cplx<T>* __restrict__ fhat, int ROWS) {
  // src is psi itself except for the first launch of a core started from another slot.
  // In place, each thread reads its own site's D values in the prologue, and the block
  // writes them only after its last barrier, from LDS: no load of src can move past a
  // store to psi.  (Without __restrict__ on src the FP64 spin-6 launch ran 6 % slower.)
  extern __shared__ __align__(16) unsigned char smem[];
  // values per thread in the tile FFTs: 16 for 13 components, so one round of TS/TEAM =
  // 16*ROWS lines covers all 13*ROWS (a butterfly's arithmetic does not depend on E)
  constexpr int EF = D > 8 ? 16 : 8;
  constexpr int TEAM = NX / EF;
  const int tid = threadIdx.x;
  const int TS = blockDim.x;  // sites of the tile = threads (the launch has ROWS*NX threads)
  const int inst = tb.active[blockIdx.y];
  const InstParams& P = tb.params[inst];
  const double dt = prog.dt_override > 0.0 ? prog.dt_override : P.dt;
  const bool itp = prog.itp != 0;
  const int st = prog.stages;
  const bool has_aux = (st & (XS_LOAD_PHI | XS_WRITE_FHAT)) != 0;
  const int AL = g.spad ? 2 * NX : NX;  // DDI auxiliary line length
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem);
  cplx<T>* aux = tile + (size_t)D * TS / 8 * 9;
  cplx<double>* zf =
      reinterpret_cast<cplx<double>*>(aux + (has_aux ? 2 * ROWS * AL / 8 * 9 : 0));
  double* rsh = reinterpret_cast<double*>(zf + kMaxOps * D);
  cplx<double>* tws = reinterpret_cast<cplx<double>*>(rsh + 16);  // NX twiddles
  const cplx<double>* twg = tw_table<NX>(tb);
  const int row0 = blockIdx.x * ROWS;
  const size_t ibase = (size_t)inst * g.inst_stride + (size_t)row0 * NX;

  // Prologue: the D tile loads of this thread's site, one twiddle and (with the DDI
  // potential stage) the spectral bins of this thread's AL/NX <= 2 line entries are all
  // issued before the first wait; TS >= NX threads cover the twiddle table.
  const size_t fb = (size_t)inst * g.fhat_inst_stride;
  const int nph = (st & XS_LOAD_PHI) ? AL / NX : 0;
  {
    const cplx<double> twv = twg[min(tid, NX - 1)];
    cplx<T> ld[D];
    const cplx<T>* pb = src + ibase;  // psi itself, or the slot a core starts from
#pragma unroll
    for (int c = 0; c < D; ++c) ld[c] = *lane_elem(pb + (size_t)c * g.nsites, tid);
    cplx<T> ph[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (k < nph) ddi_phi_fetch<T, true>(g, fhat, fb, row0, AL, tid + k * TS, ph[k]);
    tws[min(tid, NX - 1)] = twv;
#pragma unroll
    for (int c = 0; c < D; ++c) tile[p8(c * TS + tid)] = ld[c];
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (k < nph) ddi_phi_place<T>(AL, tid + k * TS, ph[k], aux);
  }
  // op list in LDS, copied with compile-time indices (the by-value kernel argument is
  // never indexed at run time), then the FP64 Zeeman factors of its diagonal steps
  __shared__ Op s_ops[kMaxOps];
#pragma unroll
  for (int i = 0; i < kMaxOps; ++i)
    if (tid == (i < TS ? i : 0)) s_ops[i] = prog.A[i];
  __syncthreads();
  for (int idx = tid; idx < kMaxOps * D; idx += TS) {
    const int i = idx / D, c = idx % D;
    const Op op = s_ops[i];
    if (i < prog.nA && op.kind == OP_DIAG) {
      const double dtf = op.frac * dt;
      if (itp) zf[idx] = mk(exp(-(P.zee[c] - P.zee_shift) * dtf), 0.0);
      else zf[idx] = cis_neg<double>(P.zee[c] * dtf);
    }
  }
  __syncthreads();

  // FFT of the D*ROWS tile lines in rounds of TS/TEAM = (EF)*ROWS lines
  auto tile_fft = [&](auto invc) __attribute__((always_inline)) {
    constexpr bool INV = decltype(invc)::value;
    const int lpr = TS / TEAM, l = tid / TEAM, t = tid - l * TEAM;
    for (int l0 = 0; l0 < D * ROWS; l0 += lpr) {
      const int ln = l0 + l;
      const bool act = ln < D * ROWS;
      fft_team<T, NX, INV, EF, true>(tile + p8((act ? ln : 0) * NX), 1, t, act, tws);
    }
  };
  if (st & XS_INV_FFT) tile_fft(std::true_type{});
  if (st & XS_LOAD_PHI) {  // lines placed in the prologue (the op-list barrier follows)
    ddi_lines_fft<T, NX, true>(tb, aux, ROWS, AL, tid, TS, tws);
    // max |Phi| for the imaginary-time overflow guard (ground_state.jl:4-29), taken before
    // the site program (a following spin-density stage reuses the lines); the guard reads
    // the unpadded Phi buffers, which the padded convolution does not write
    double phimax = 0.0;
    if (!g.spad) {
      const int r = tid / NX, x = tid - r * NX;
      const cplx<T> a = aux[p8((2 * r) * AL + x)], b = aux[p8((2 * r + 1) * AL + x)];
      phimax = fmax(phimax, fmax(fabs((double)a.x), fmax(fabs((double)a.y), fabs((double)b.x))));
    }
    const double m = block_max(phimax, rsh);
    if (tid == 0) *partial_ptr(tb, RS_PHIMAX, inst, blockIdx.x) = m;
  }

  // The site program, one spinor per thread.  Only the spinor is live across the
  // operators: Phi is read from LDS inside the rotation (before a spin-density stage of
  // the same launch overwrites this thread's own slots), and the norm partial is taken
  // afterwards from LDS, as in the x pass.
  int nmv = 0, nrot = 0;
  {
    cplx<T> sp[D];
#pragma unroll
    for (int c = 0; c < D; ++c) sp[c] = tile[p8(c * TS + tid)];
    const long long site = (long long)row0 * NX + tid;
    for (int i = 0; i < prog.nA; ++i) {
      const Op op = s_ops[i];
      const double dtf = op.frac * dt;
      if (op.kind == OP_DIAG) {
        op_diag<T, D>(sp, trap_site<true>(g, tb, inst, site), P, dtf, itp, &zf[i * D]);
      } else if (op.kind == OP_SPINMIX) {
        nrot += op_spin_mixing<T, D>(sp, P, dtf, itp);
      } else if (op.kind == OP_DDI_ROT) {
        const int r = tid / NX, x = tid - r * NX;
        const cplx<T> a = aux[p8((2 * r) * AL + x)], b = aux[p8((2 * r + 1) * AL + x)];
        nrot += op_euler<T, D>(sp, (double)a.x, (double)a.y, (double)b.x, dtf, itp);
      } else if (op.kind == OP_NEMATIC) {
        op_nematic<T, D>(sp, P.c2, dtf, itp);
      } else if (op.kind == OP_LOSS) {
        op_loss<T, D>(sp, P, dtf);
      } else if (op.kind == OP_RAMAN) {
        if (P.has_raman) {
          double phx, phy, phz;
          raman_phi<true>(g, tb, P, site, phx, phy, phz);
          nrot += op_euler<T, D>(sp, phx, phy, phz, dtf, itp);
        }
      } else if (op.kind == OP_TENSOR) {
        if constexpr (D >= 3 && sizeof(T) == 8) {
          // FP64: the Taylor sum lives in this thread's own tile slots (written back
          // below), which keeps the step within 256 VGPRs (2 waves/SIMD)
          auto at = [TS, tid](int c) __attribute__((always_inline)) { return p8(c * TS + tid); };
          LdsAcc<T, decltype(at)> acc{tile, at};
          nmv += op_tensor<T, D>(sp, P, dtf, itp, acc, tb.kfold + (size_t)inst * tens::kKfoldMax);
        } else if constexpr (D >= 3) {
          nmv += op_tensor<T, D>(sp, P, dtf, itp, tb.kfold + (size_t)inst * tens::kKfoldMax);
        }
      } else if (op.kind == OP_SCALE) {
        const double sc = 1.0 / sqrt(tb.red[(size_t)inst * RS_COUNT + RS_NORM] * g.dV);
#pragma unroll
        for (int c = 0; c < D; ++c) sp[c] = scale64(sp[c], sc);
      }
    }
    if (st & XS_WRITE_FHAT) {  // spin density into this thread's own slots of the lines
      T fx, fy, fz;
      spin_vec<T, D>(sp, fx, fy, fz);
      const int r = tid / NX, x = tid - r * NX;
      aux[p8((2 * r) * AL + x)] = mk(fx, fy);
      aux[p8((2 * r + 1) * AL + x)] = mk(fz, (T)0);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) tile[p8(c * TS + tid)] = sp[c];
  }
  double norm_acc = 0.0;
  if (st & XS_NORM) {  // own elements of the tile
#pragma unroll
    for (int c = 0; c < D; ++c) norm_acc += (double)abs2(tile[p8(c * TS + tid)]);
  }

  if (st & XS_WRITE_FHAT) {
    if (AL > NX) ddi_lines_zero_upper<T, NX>(aux, ROWS, AL, tid, TS);
    __syncthreads();
    ddi_lines_fft<T, NX, false>(tb, aux, ROWS, AL, tid, TS, tws);
    ddi_lines_out<T, true>(g, fhat, fb, row0, ROWS, AL, aux, tid, TS);
  }
  if (st & XS_FWD_FFT) {
    __syncthreads();
    tile_fft(std::false_type{});
  }
  // without the forward FFT these are the thread's own writes; the FFT ends on a barrier.
  // The addresses are formed again here from an opaque base and lane index (kept from the
  // prologue, the 13 FP64 addresses would spill across the site program).
  {
    cplx<T>* pb = psi + ibase;
    int lane = tid;
    asm volatile("" : "+s"(pb), "+v"(lane));
#pragma unroll
    for (int c = 0; c < D; ++c) *lane_elem(pb + (size_t)c * g.nsites, lane) = tile[p8(c * TS + tid)];
  }

  // per-block partial (every thread takes part)
  if (st & XS_NORM) {
    const double v = block_sum(norm_acc, rsh);
    if (tid == 0) *partial_ptr(tb, RS_NORM, inst, blockIdx.x) = v;
  }
  if (tb.work) {  // profiling: executed tensor matvecs and spin rotations
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

static size_t xsite_lds(int D, int rows, int NX, size_t csz, bool has_aux, int AL) {
  return csz * ((size_t)D * rows * NX + (has_aux ? 2 * (size_t)rows * AL : 0)) / 8 * 9 +
         sizeof(cplx<double>) * kMaxOps * D + 16 * sizeof(double) + sizeof(cplx<double>) * NX;
}

bool xsite_ok(const Geom& g) {
  if (g.D != 5 && g.D != 13) return false;
  if (g.ndim < 2) return false;
  const int nx = g.n[0];
  if (nx < 16 || nx > 256) return false;
  // power-of-two sides on every axis: site coordinates and spectra rows by shifts
  for (int d = 0; d < 3; ++d)
    if ((g.n[d] & (g.n[d] - 1)) != 0) return false;
  const int ts = xpass_rows_g(g) * nx;
  // one thread per site (tiles below one wave, from short lines or the tile autotuner's
  // choice, leave lanes idle but keep the x pass's block partition)
  return ts >= 16 && ts <= kXsiteMaxThreads;
}

template <typename T, int NX, int D>
static hipError_t xsite_n(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                          const void* src, void* psi, void* fhat, int n_active) {
  const int rows = xpass_rows_g(g);
  const int threads = rows * NX;
  if (threads < 16 || threads > kXsiteMaxThreads) return hipErrorInvalidConfiguration;
  const bool has_aux = (prog.stages & (XS_LOAD_PHI | XS_WRITE_FHAT)) != 0;
  const size_t lds = xsite_lds(D, rows, NX, sizeof(cplx<T>), has_aux, g.spad ? 2 * NX : NX);
  if (lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void*)xsite_kernel<T, NX, D>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const dim3 grid(g.nrows / rows, n_active);
  hipLaunchKernelGGL((xsite_kernel<T, NX, D>), grid, dim3(threads), lds, st, g, tb, prog,
                     (const cplx<T>*)src, (cplx<T>*)psi, (cplx<T>*)fhat, rows);
  return hipGetLastError();
}

template <typename T, int D>
static hipError_t xsite_d(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                          const void* src, void* psi, void* fhat, int n_active) {
  switch (g.n[0]) {
    case 16: return xsite_n<T, 16, D>(st, g, tb, prog, src, psi, fhat, n_active);
    case 32: return xsite_n<T, 32, D>(st, g, tb, prog, src, psi, fhat, n_active);
    case 64: return xsite_n<T, 64, D>(st, g, tb, prog, src, psi, fhat, n_active);
    case 128: return xsite_n<T, 128, D>(st, g, tb, prog, src, psi, fhat, n_active);
    case 256: return xsite_n<T, 256, D>(st, g, tb, prog, src, psi, fhat, n_active);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_xsite(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                        const void* src, void* psi, void* fhat, int n_active) {
  // the stages this kernel implements; the schedule never asks it for others
  constexpr int kStages = XS_INV_FFT | XS_LOAD_PHI | XS_OPS_A | XS_WRITE_FHAT | XS_NORM | XS_FWD_FFT;
  if ((prog.stages & ~kStages) != 0 || prog.nB != 0 || prog.nA > kMaxOps) return hipErrorInvalidValue;
  switch (g.D) {
    case 5: return xsite_d<T, 5>(st, g, tb, prog, src, psi, fhat, n_active);
    case 13: return xsite_d<T, 13>(st, g, tb, prog, src, psi, fhat, n_active);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_xsite<float>(hipStream_t, const Geom&, const DevTables&, const XProgram&,
                                        const void*, void*, void*, int);
template hipError_t launch_xsite<double>(hipStream_t, const Geom&, const DevTables&,
                                         const XProgram&, const void*, void*, void*, int);

}  // namespace bec
