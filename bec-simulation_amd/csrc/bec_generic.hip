// Line passes for grid sides without a built radix plan.
//
// Any even side (up to what one block's LDS holds) runs through these two kernels, which
// keep the stages of the x pass (bec_kernels.hip) and the modes of the y/z passes
// (bec_strided.hip) with the same memory layout, partial-sum slots and reduction blocks,
// so the schedule, the reductions and the lane-per-site kernel are shared with the built
// lengths.  What differs is the FFT: a Stockham transform whose radices (4 while four
// divides the remaining length, then 2, then every odd prime factor in increasing order)
// are found at run time.  Each output of a radix-R pass is one thread's R-term sum
//   out[k] = sum_t in[j + t N/R] * w^(t * (k mod Ns R) * N/(Ns R)),   w = exp(-+2 pi i / N)
// (j = (k div Ns R) Ns + k mod Ns, Ns = product of the radices before), read from one LDS
// region and written to another, with the powers of w taken from an N-entry FP64 table
// (exact integer exponents mod N).  Sums are formed in FP64 for both storage precisions,
// so a pass costs N * R complex multiply-adds per line: these kernels trade the built
// lengths' register-resident radix-8 butterflies for generality.  Lengths whose large prime
// factor would make that sum dominate run Bluestein's algorithm instead (gen_fft_blue,
// chosen per length by gen_blue_m).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "bec_device.h"
#include "bec_kernels.h"
#include "bec_passes.h"

namespace bec {

constexpr int kGenThreads = 256;

// Division by a block-uniform runtime divisor d >= 1 for 0 <= x < 2^22: a float reciprocal
// gives the quotient to within one, and one remainder test fixes it (exact; about a
// quarter of the instructions of a general 32-bit division).
struct FastDiv {
  int d;
  float r;
  __device__ __forceinline__ explicit FastDiv(int dv) : d(dv), r(1.0f / (float)dv) {}
  __device__ __forceinline__ int div(int x) const {
    int q = (int)((float)x * r);
    const int rem = x - q * d;
    q += rem >= d ? 1 : 0;
    q -= rem < 0 ? 1 : 0;
    return q;
  }
};

// Twiddle table exp(-2 pi i m / N), m < N, of a runtime line length (the workspace uploads
// one for every side and zero-padded spectral length of its grid).
// (The launchers check on the host that the table exists.)
__device__ __forceinline__ const cplx<double>* gen_tw(const DevTables& tb, int N) {
  const void* p = nullptr;
#pragma unroll
  for (int i = 0; i < kGenTabs; ++i)
    if (tb.gtw_n[i] == N) p = tb.gtw[i];
  return reinterpret_cast<const cplx<double>*>(p);
}
static bool has_gen_tw(const DevTables& tb, int N) {
  for (int i = 0; i < kGenTabs; ++i)
    if (tb.gtw_n[i] == N && tb.gtw[i]) return true;
  return false;
}

// Bluestein tables of a line length (M = 0: the length has none and runs mixed-radix passes).
struct GenBlue {
  int M;
  const cplx<double>* c;   // chirp exp(i pi n^2 / N), n < N
  const cplx<double>* ch;  // M-point transform of the wrapped chirp, divided by M
  const cplx<double>* tm;  // exp(-2 pi i m / M), m < M
};
__device__ __forceinline__ GenBlue gen_blue(const DevTables& tb, int N) {
  GenBlue b{0, nullptr, nullptr, nullptr};
#pragma unroll
  for (int i = 0; i < kGenTabs; ++i) {
    if (tb.gtw_n[i] == N && tb.gbl_m[i] > 0) {
      b.M = tb.gbl_m[i];
      b.c = reinterpret_cast<const cplx<double>*>(tb.gbl[i]);
      b.ch = b.c + N;
      b.tm = b.ch + b.M;
    }
  }
  return b;
}

// One radix-R Stockham pass over nl lines (line l at src + l * P), into dst: output
//   out[k] = sum_t in[j + t N/R] * w^(t * (k mod Ns R) * N/(Ns R)),   w = exp(-+2 pi i / N)
// with j = (k div Ns R) Ns + k mod Ns (Ns = product of the radices before).
//
// Radices 2, 3, 4, 5 and 7 run as butterflies: one thread takes the R outputs
// k = jh Ns R + jl + q Ns (q < R) that share the inputs in[j + t N/R], j = jh Ns + jl; it
// multiplies input t by w^(t jl N/(Ns R)) (an exponent below N) and takes the R-point DFT of
// the products in registers (constants cos/sin(2 pi m / R)).  Larger prime radices keep one
// R-term sum per output.  Sums in FP64 for both storage precisions.
template <int R>
struct SmallDft;  // e^(-2 pi i m / R), m < R
template <> struct SmallDft<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, -0.8660254037844386, 0.8660254037844386};
};
template <> struct SmallDft<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475,
                                  -0.8090169943749475, 0.30901699437494745};
  static constexpr double s[5] = {0.0, -0.9510565162951535, -0.5877852522924731,
                                  0.5877852522924731, 0.9510565162951535};
};
template <> struct SmallDft<7> {
  static constexpr double c[7] = {1.0, 0.6234898018587336, -0.22252093395631434,
                                  -0.9009688679024191, -0.9009688679024191,
                                  -0.22252093395631434, 0.6234898018587336};
  static constexpr double s[7] = {0.0, -0.7818314824680298, -0.9749279121818236,
                                  -0.43388373911755823, 0.43388373911755823,
                                  0.9749279121818236, 0.7818314824680298};
};

template <typename T, int R>
__device__ __forceinline__ void gen_butterflies(const cplx<T>* __restrict__ src,
                                                cplx<T>* __restrict__ dst, int P, int nl, int N,
                                                int Ns, const cplx<double>* __restrict__ tw,
                                                bool inv, int tid, int nthr) {
  const int NB = N / R, step = N / (Ns * R);
  const FastDiv fnb(NB), fns(Ns);
  for (int b = tid; b < nl * NB; b += nthr) {
    const int l = fnb.div(b), j = b - l * NB;
    const int jh = fns.div(j), jl = j - jh * Ns;
    const cplx<T>* in = src + (size_t)l * P + j;
    double vx[R], vy[R];
    vx[0] = (double)in[0].x;
    vy[0] = (double)in[0].y;
    const int e1 = jl * step;
#pragma unroll
    for (int t = 1; t < R; ++t) {
      const cplx<double> w = tw[t * e1];
      const double wy = inv ? -w.y : w.y;
      const cplx<T> v = in[t * NB];
      vx[t] = (double)v.x * w.x - (double)v.y * wy;
      vy[t] = (double)v.x * wy + (double)v.y * w.x;
    }
    cplx<T>* out = dst + (size_t)l * P + jh * Ns * R + jl;
    if constexpr (R == 2) {
      out[0] = mk((T)(vx[0] + vx[1]), (T)(vy[0] + vy[1]));
      out[Ns] = mk((T)(vx[0] - vx[1]), (T)(vy[0] - vy[1]));
    } else if constexpr (R == 4) {
      // X_q = (a0 + a2) + (-+i)^q (a1 + a3) ... with -i for the forward transform
      const double s0x = vx[0] + vx[2], s0y = vy[0] + vy[2];
      const double d0x = vx[0] - vx[2], d0y = vy[0] - vy[2];
      const double s1x = vx[1] + vx[3], s1y = vy[1] + vy[3];
      const double d1x = vx[1] - vx[3], d1y = vy[1] - vy[3];
      // (-i) d1 forward, (+i) d1 inverse
      const double rx = inv ? -d1y : d1y, ry = inv ? d1x : -d1x;
      out[0] = mk((T)(s0x + s1x), (T)(s0y + s1y));
      out[Ns] = mk((T)(d0x + rx), (T)(d0y + ry));
      out[2 * Ns] = mk((T)(s0x - s1x), (T)(s0y - s1y));
      out[3 * Ns] = mk((T)(d0x - rx), (T)(d0y - ry));
    } else {
      using W = SmallDft<R>;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        double ax = vx[0], ay = vy[0];
#pragma unroll
        for (int t = 1; t < R; ++t) {
          const int m = (t * q) % R;
          const double c = W::c[m], sn = inv ? -W::s[m] : W::s[m];
          ax += vx[t] * c - vy[t] * sn;
          ay += vx[t] * sn + vy[t] * c;
        }
        out[q * Ns] = mk((T)ax, (T)ay);
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ void gen_stage(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst,
                                          int P, int nl, int N, int R, int Ns,
                                          const cplx<double>* __restrict__ tw, bool inv, int tid,
                                          int nthr) {
  switch (R) {
    case 2: gen_butterflies<T, 2>(src, dst, P, nl, N, Ns, tw, inv, tid, nthr); return;
    case 3: gen_butterflies<T, 3>(src, dst, P, nl, N, Ns, tw, inv, tid, nthr); return;
    case 4: gen_butterflies<T, 4>(src, dst, P, nl, N, Ns, tw, inv, tid, nthr); return;
    case 5: gen_butterflies<T, 5>(src, dst, P, nl, N, Ns, tw, inv, tid, nthr); return;
    case 7: gen_butterflies<T, 7>(src, dst, P, nl, N, Ns, tw, inv, tid, nthr); return;
    default: break;
  }
  const int blk = Ns * R, step = N / blk, stride = N / R;
  const FastDiv fn(N), fblk(blk), fns(Ns);
  for (int o = tid; o < nl * N; o += nthr) {
    const int l = fn.div(o), k = o - l * N;
    const int jh = fblk.div(k), rem = k - jh * blk;
    const int jl = rem - fns.div(rem) * Ns;
    const cplx<T>* in = src + (size_t)l * P + jh * Ns + jl;
    const int s = rem * step;
    double ax = (double)in[0].x, ay = (double)in[0].y;
    int idx = 0;
    for (int t = 1; t < R; ++t) {
      idx += s;
      if (idx >= N) idx -= N;
      const cplx<double> w = tw[idx];
      const double wy = inv ? -w.y : w.y;
      const cplx<T> v = in[t * stride];
      ax += (double)v.x * w.x - (double)v.y * wy;
      ay += (double)v.x * wy + (double)v.y * w.x;
    }
    dst[(size_t)l * P + k] = mk((T)ax, (T)ay);
  }
}

// Bluestein's algorithm for lengths with a large prime factor, whose radix-p pass would cost
// N p complex multiply-adds per line: with nk = (n^2 + k^2 - (k - n)^2) / 2 and
// c[n] = exp(i pi n^2 / N), the forward DFT is X[k] = conj(c[k]) sum_n x[n] conj(c[n]) c[k - n],
// a cyclic convolution with the chirp once both are zero-padded to M >= 2N - 1 points
// (M a power of two): transform, multiply by the chirp's transform, transform back.  The
// inverse DFT is the same with c and conj(c) exchanged (the wrapped chirp is symmetric, so
// its transform is conjugated).  Lines in groups that fit the scratch region (two M-point
// work lines each); every product in FP64.
template <typename T>
__device__ __forceinline__ void gen_fft_blue(cplx<T>* buf, cplx<T>* scr, int nl, int N, int P,
                                             int cap_elems, const GenBlue& bl, bool inv, int tid,
                                             int nthr) {
  // work lines in FP64 for both storage precisions (the M-point values grow to sqrt(M)
  // times the input, which FP32 storage between passes would round at that scale), from
  // the first 16-byte boundary of the scratch region
  const int M = bl.M, msh = __ffs(M) - 1;
  cplx<double>* wk = reinterpret_cast<cplx<double>*>(
      (reinterpret_cast<uintptr_t>(scr) + 15) & ~static_cast<uintptr_t>(15));
  const long long bytes = (long long)cap_elems * (long long)sizeof(cplx<T>) - 16;
  const int gcap = max(1, (int)(bytes / (2LL * M * (long long)sizeof(cplx<double>))));
  auto cmul = [](cplx<double> x, cplx<double> w) __attribute__((always_inline)) {
    return mk(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
  };
  for (int l0 = 0; l0 < nl; l0 += gcap) {
    const int ng = min(gcap, nl - l0);
    cplx<double>* a = wk;
    cplx<double>* b = wk + (size_t)ng * M;
    for (int e = tid; e < ng * M; e += nthr) {
      const int l = e >> msh, n = e & (M - 1);
      cplx<double> v = mk(0.0, 0.0);
      if (n < N) {
        const cplx<double> c = bl.c[n];
        const cplx<T> x = buf[(size_t)(l0 + l) * P + n];
        v = cmul(mk((double)x.x, (double)x.y), mk(c.x, inv ? c.y : -c.y));
      }
      a[e] = v;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      int rem = M, Ns = 1;
      while (rem > 1) {
        const int R = (rem & 3) == 0 ? 4 : 2;
        gen_stage<double>(a, b, M, ng, M, R, Ns, bl.tm, pass == 1, tid, nthr);
        __syncthreads();
        cplx<double>* t = a;
        a = b;
        b = t;
        Ns *= R;
        rem /= R;
      }
      if (pass == 0) {
        for (int e = tid; e < ng * M; e += nthr) {
          const cplx<double> h = bl.ch[e & (M - 1)];
          a[e] = cmul(a[e], mk(h.x, inv ? -h.y : h.y));
        }
        __syncthreads();
      }
    }
    const FastDiv fn(N);
    for (int e = tid; e < ng * N; e += nthr) {
      const int l = fn.div(e), k = e - l * N;
      const cplx<double> c = bl.c[k];
      const cplx<double> y = cmul(a[((size_t)l << msh) + k], mk(c.x, inv ? c.y : -c.y));
      buf[(size_t)(l0 + l) * P + k] = mk((T)y.x, (T)y.y);
    }
    __syncthreads();
  }
}

// Unnormalised FFT (forward: exp(-2 pi i jk/N)) of nl lines of length N at buf + l * P, in
// place, through a scratch region of `cap` lines of the same pitch (groups of cap lines);
// lengths with Bluestein tables (bl.M > 0) go through gen_fft_blue.
// Every thread of the block calls it; it ends with the block synchronised.
template <typename T>
__device__ __forceinline__ void gen_fft(cplx<T>* buf, cplx<T>* scr, int nl, int N, int P, int cap,
                        const cplx<double>* tw, bool inv, int tid, int nthr, const GenBlue& bl) {
  if (bl.M > 0) {
    gen_fft_blue<T>(buf, scr, nl, N, P, cap * P, bl, inv, tid, nthr);
    return;
  }
  for (int l0 = 0; l0 < nl; l0 += cap) {
    const int ng = min(cap, nl - l0);
    cplx<T>* a = buf + (size_t)l0 * P;
    bool in_scr = false;
    int rem = N, Ns = 1;
    while (rem > 1) {
      int R;
      if ((rem & 3) == 0) R = 4;
      else if ((rem & 1) == 0) R = 2;
      else {
        R = 3;
        while (rem % R) R += 2;
      }
      gen_stage<T>(in_scr ? scr : a, in_scr ? a : scr, P, ng, N, R, Ns, tw, inv, tid, nthr);
      __syncthreads();
      in_scr = !in_scr;
      Ns *= R;
      rem /= R;
    }
    if (in_scr) {
      const FastDiv fn(N);
      for (int o = tid; o < ng * N; o += nthr) {
        const int l = fn.div(o), k = o - l * N;
        a[(size_t)l * P + k] = scr[(size_t)l * P + k];
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// x pass (stages as xpass_kernel): ROWS whole x rows of every component in LDS, line
// (c, r) at tile + (c ROWS + r) (NX + 1); DDI auxiliary lines (2 per row, length sn[0]) at
// aux + l (sn[0] + 1).  The odd pitches keep column-wise LDS accesses off a single bank.
// ---------------------------------------------------------------------------
template <typename T, int D>
__global__ void __launch_bounds__(kGenThreads)
xgen_kernel(Geom g, DevTables tb, XProgram prog, const cplx<T>* __restrict__ src,
            cplx<T>* __restrict__ dst, cplx<T>* __restrict__ fhat, T* __restrict__ phi, int ROWS,
            int cap_elems) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NX = g.n[0];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int inst = tb.active[blockIdx.y];
  const InstParams& P = tb.params[inst];
  const double dt = prog.dt_override > 0.0 ? prog.dt_override : P.dt;
  const bool itp = prog.itp != 0;
  const int st = prog.stages;
  const bool has_aux = (st & (XS_LOAD_PHI | XS_WRITE_FHAT)) != 0;
  const int AL = g.sn[0];
  const int LP = NX + 1, ALP = AL + 1;
  const int TS = ROWS * NX;
  const int NL = D * ROWS;
  cplx<double>* twx = reinterpret_cast<cplx<double>*>(smem);
  cplx<double>* twa = AL != NX ? twx + NX : twx;
  cplx<double>* zf = twx + NX + (AL != NX ? AL : 0);
  double* rsh = reinterpret_cast<double*>(zf + 2 * kMaxOps * D);
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(rsh + 16);
  cplx<T>* aux = tile + (size_t)NL * LP;
  cplx<T>* scr = aux + (has_aux ? (size_t)2 * ROWS * ALP : 0);
  const int capx = max(1, cap_elems / LP), capa = max(1, cap_elems / ALP);
  const GenBlue blx = gen_blue(tb, NX), bla = gen_blue(tb, AL);

  const int row0 = blockIdx.x * ROWS;
  const size_t ibase = (size_t)inst * g.inst_stride + (size_t)row0 * NX;
  const FastDiv fNX(NX), fRows(ROWS);
  {
    const cplx<double>* t = gen_tw(tb, NX);
    for (int i = tid; i < NX; i += nthr) twx[i] = t[i];
    if (AL != NX) {
      const cplx<double>* ta = gen_tw(tb, AL);
      for (int i = tid; i < AL; i += nthr) twa[i] = ta[i];
    }
  }
  if (!(st & XS_NO_LOAD)) {
    for (int e = tid; e < NL * NX; e += nthr) {
      const int l = fNX.div(e), x = e - l * NX;
      const int c = fRows.div(l), r = l - c * ROWS;
      tile[(size_t)l * LP + x] = src[ibase + (size_t)c * g.nsites + (size_t)r * NX + x];
    }
  }
  __shared__ Op s_ops[2 * kMaxOps];
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < kMaxOps; ++i) {
      s_ops[i] = prog.A[i];
      s_ops[kMaxOps + i] = prog.B[i];
    }
  }
  __syncthreads();
  for (int idx = tid; idx < 2 * kMaxOps * D; idx += nthr) {
    const int i = idx / D, c = idx % D;
    const bool used = i < kMaxOps ? i < prog.nA : (i - kMaxOps) < prog.nB;
    const Op op = s_ops[i];
    if (used && op.kind == OP_DIAG) {
      const double dtf = op.frac * dt;
      if (itp) zf[idx] = mk(exp(-(P.zee[c] - P.zee_shift) * dtf), 0.0);
      else zf[idx] = cis_neg<double>(P.zee[c] * dtf);
    }
  }
  __syncthreads();
  double norm_acc = 0.0, ke_acc = 0.0, phimax = 0.0;

  if (st & XS_INV_FFT) gen_fft<T>(tile, scr, NL, NX, LP, capx, twx, true, tid, nthr, blx);

  const double sc = 1.0 / sqrt(tb.red[(size_t)inst * RS_COUNT + RS_NORM] * g.dV);

  if (st & XS_LOAD_PHI) {
    // C2R input (reference irfft semantics, ddi.jl:108-136), two-for-one as in ddi_phi_place
    const size_t fb = (size_t)inst * g.fhat_inst_stride;
    for (int e = tid; e < ROWS * AL; e += nthr) {
      const int r = e / AL, x = e - r * AL;
      const size_t o = fb + (size_t)spec_row(g, row0 + r) * g.nxh_pitch;
      const int kk = x <= AL / 2 ? x : AL - x;
      cplx<T> yx = fhat[o + kk], yy = fhat[o + g.fhat_field_stride + kk],
              yz = fhat[o + 2 * g.fhat_field_stride + kk];
      if (kk == 0 || kk == AL / 2) {
        yx.y = 0; yy.y = 0; yz.y = 0;
      } else if (x > AL / 2) {
        yx = conjc(yx); yy = conjc(yy); yz = conjc(yz);
      }
      aux[(size_t)(2 * r) * ALP + x] = mk(yx.x - yy.y, yx.y + yy.x);
      aux[(size_t)(2 * r + 1) * ALP + x] = yz;
    }
    __syncthreads();
    gen_fft<T>(aux, scr, 2 * ROWS, AL, ALP, capa, twa, true, tid, nthr, bla);
    if (!g.spad) {  // max |Phi| (ground_state.jl:4-29), unpadded convolutions only
      for (int s = tid; s < TS; s += nthr) {
        const int r = fNX.div(s), x = s - r * NX;
        const cplx<T> a = aux[(size_t)(2 * r) * ALP + x], b = aux[(size_t)(2 * r + 1) * ALP + x];
        phimax = fmax(phimax, fmax(fabs((double)a.x), fmax(fabs((double)a.y), fabs((double)b.x))));
      }
    }
    if (st & XS_STORE_PHI) {
      T* ph = phi + (size_t)inst * 3 * g.nsites + (size_t)row0 * NX;
      for (int s = tid; s < TS; s += nthr) {
        const int r = fNX.div(s), x = s - r * NX;
        const cplx<T> a = aux[(size_t)(2 * r) * ALP + x], b = aux[(size_t)(2 * r + 1) * ALP + x];
        ph[s] = a.x;
        ph[g.nsites + s] = a.y;
        ph[2 * g.nsites + s] = b.x;
      }
    }
  }

  // light per-site operators (the schedule runs the heavy ones in the lane-per-site kernel)
  auto run_ops = [&](const Op* ops, int nops, int zbase) {
    for (int i = 0; i < nops; ++i) {
      const Op op = ops[i];
      const double dtf = op.frac * dt;
      if (op.kind == OP_SCALE) {
        for (int e = tid; e < NL * NX; e += nthr) {
          const int l = fNX.div(e), x = e - l * NX;
          cplx<T>& v = tile[(size_t)l * LP + x];
          v = scale64(v, sc);
        }
      } else {
        for (int s = tid; s < TS; s += nthr) {
          const int r = fNX.div(s), x = s - r * NX;
          cplx<T> sp[D];
#pragma unroll
          for (int c = 0; c < D; ++c) sp[c] = tile[(size_t)(c * ROWS + r) * LP + x];
          if (op.kind == OP_DIAG) {
            const double Vt = trap_site(g, tb, inst, (long long)(row0 + r) * NX + x);
            op_diag<T, D>(sp, Vt, P, dtf, itp, &zf[(zbase + i) * D]);
          } else if (op.kind == OP_SPINMIX) {
            if constexpr (D == 3) op_spin_mixing<T, D>(sp, P, dtf, itp);
          } else if (op.kind == OP_NEMATIC) {
            if constexpr (D < 13) op_nematic<T, D>(sp, P.c2, dtf, itp);
          } else if (op.kind == OP_LOSS) {
            if (P.has_loss) op_loss<T, D>(sp, P, dtf);
          }
#pragma unroll
          for (int c = 0; c < D; ++c) tile[(size_t)(c * ROWS + r) * LP + x] = sp[c];
        }
      }
      __syncthreads();  // the scale loop's element ownership differs from the site loops
    }
  };

  if (st & XS_OPS_A) run_ops(s_ops, prog.nA, 0);

  if (st & XS_KIN_1D) {
    gen_fft<T>(tile, scr, NL, NX, LP, capx, twx, false, tid, nthr, blx);
    const double kdt = prog.kin_frac * dt;
    for (int e = tid; e < NL * NX; e += nthr) {
      const int l = fNX.div(e), x = e - l * NX;
      const double k = tb.ktab[x];
      cplx<double> ph;
      if (prog.deriv) {  // fft_utils.jl:4-19 along x: i k_x / N on component 0
        ph = mk(0.0, l < ROWS ? k * g.inv_ntot : 0.0);
      } else {
        const double a = 0.5 * (k * k) * kdt;
        ph = itp ? mk(exp(-a), 0.0) : cis_neg<double>(a);
        ph = ph * g.inv_ntot;
      }
      cplx<T>& v = tile[(size_t)l * LP + x];
      v = mul64(v, ph);
    }
    __syncthreads();
    gen_fft<T>(tile, scr, NL, NX, LP, capx, twx, true, tid, nthr, blx);
  }

  if (st & XS_OPS_B) run_ops(s_ops + kMaxOps, prog.nB, kMaxOps);

  if (st & XS_WRITE_FHAT) {
    for (int s = tid; s < TS; s += nthr) {
      const int r = fNX.div(s), x = s - r * NX;
      cplx<T> sp[D];
#pragma unroll
      for (int c = 0; c < D; ++c) sp[c] = tile[(size_t)(c * ROWS + r) * LP + x];
      T fx, fy, fz;
      spin_vec<T, D>(sp, fx, fy, fz);
      aux[(size_t)(2 * r) * ALP + x] = mk(fx, fy);
      aux[(size_t)(2 * r + 1) * ALP + x] = mk(fz, (T)0);
    }
    if (AL > NX) {  // zero padding of the spin-density lines
      const int W = AL - NX;
      for (int e = tid; e < 2 * ROWS * W; e += nthr) {
        const int l = e / W, x = NX + (e - l * W);
        aux[(size_t)l * ALP + x] = mk((T)0, (T)0);
      }
    }
    __syncthreads();
    gen_fft<T>(aux, scr, 2 * ROWS, AL, ALP, capa, twa, false, tid, nthr, bla);
    // R2C output: line 2r holds F_x + i F_y (split two-for-one), line 2r+1 holds F_z
    const size_t fb = (size_t)inst * g.fhat_inst_stride;
    const int NH = AL / 2 + 1;
    for (int e = tid; e < ROWS * NH; e += nthr) {
      const int r = e / NH, k = e - r * NH;
      const cplx<T> z1 = aux[(size_t)(2 * r) * ALP + k];
      const cplx<T> z1n = conjc(aux[(size_t)(2 * r) * ALP + (k == 0 ? 0 : AL - k)]);
      const cplx<T> sum = z1 + z1n, dif = z1 - z1n;
      const size_t o = fb + (size_t)spec_row(g, row0 + r) * g.nxh_pitch + k;
      fhat[o] = mk(sum.x * (T)0.5, sum.y * (T)0.5);
      fhat[o + g.fhat_field_stride] = mk(dif.y * (T)0.5, -dif.x * (T)0.5);
      fhat[o + 2 * g.fhat_field_stride] = aux[(size_t)(2 * r + 1) * ALP + k];
    }
  }

  if (st & XS_NORM) {
    for (int s = tid; s < TS; s += nthr) {
      const int r = fNX.div(s), x = s - r * NX;
#pragma unroll
      for (int c = 0; c < D; ++c) norm_acc += (double)abs2(tile[(size_t)(c * ROWS + r) * LP + x]);
    }
  }

  if (st & XS_FWD_FFT) gen_fft<T>(tile, scr, NL, NX, LP, capx, twx, false, tid, nthr, blx);

  if (st & XS_KE_1D) {
    for (int e = tid; e < NL * NX; e += nthr) {
      const int l = fNX.div(e), x = e - l * NX;
      const double k = tb.ktab[x];
      ke_acc += (k * k) * (double)abs2(tile[(size_t)l * LP + x]);
    }
  }

  if (!(st & XS_NO_STORE)) {
    for (int e = tid; e < NL * NX; e += nthr) {
      const int l = fNX.div(e), x = e - l * NX;
      const int c = fRows.div(l), r = l - c * ROWS;
      dst[ibase + (size_t)c * g.nsites + (size_t)r * NX + x] = tile[(size_t)l * LP + x];
    }
  }

  if (st & XS_LOAD_PHI) {
    const double m = block_max(phimax, rsh);
    if (tid == 0) *partial_ptr(tb, RS_PHIMAX, inst, blockIdx.x) = m;
  }
  if (st & XS_NORM) {
    const double v = block_sum(norm_acc, rsh);
    if (tid == 0) *partial_ptr(tb, RS_NORM, inst, blockIdx.x) = v;
  }
  if (st & XS_KE_1D) {
    const double v = block_sum(ke_acc, rsh);
    if (tid == 0) *partial_ptr(tb, RS_KIN, inst, blockIdx.x) = v;
  }
}

// ---------------------------------------------------------------------------
// y / z pass (modes as strided_kernel): TB contiguous x columns by the whole line, line
// (field f, column xi) at tile + (f TB + xi) (N + 1).
// ---------------------------------------------------------------------------
template <typename T, int MODE, int FIELDS>
__global__ void __launch_bounds__(kGenThreads)
sgen_kernel(Geom g, DevTables tb, cplx<T>* __restrict__ data, int axis, int TB, int cap,
            double kin_frac, double dt_override, int itp, int ddi_data) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = ddi_data ? g.sn[axis] : g.n[axis];
  const int LP = N + 1;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int inst = tb.active[blockIdx.z];
  cplx<double>* tw = reinterpret_cast<cplx<double>*>(smem);
  double* tabs = reinterpret_cast<double*>(tw + N);  // [2*(N+TB+1)]
  double* rsh = tabs + 2 * (N + TB + 1);
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(rsh + 16);
  cplx<T>* scr = tile + (size_t)FIELDS * TB * LP;
  auto at = [&](int f, int xi, int e) -> cplx<T>& { return tile[(size_t)(f * TB + xi) * LP + e]; };
  const GenBlue bl = gen_blue(tb, N);
  const int tbs = __ffs(TB) - 1;  // TB is a power of two
  const FastDiv fN(N);

  const int pitch = ddi_data ? g.nxh_pitch : g.n[0];
  const int valid = ddi_data ? g.nxh : g.n[0];
  const int n1 = ddi_data ? g.sn[1] : g.n[1];
  const long long fstride = ddi_data ? g.fhat_field_stride : g.nsites;
  const long long istride = ddi_data ? g.fhat_inst_stride : g.inst_stride;
  const long long estride = axis == 1 ? (long long)pitch : (long long)pitch * n1;
  const long long ostride = axis == 1 ? (long long)pitch * n1 : (long long)pitch;
  const bool padded = ddi_data && g.spad;
  const int n_in = (padded && MODE != SM_INV) ? g.n[axis] : N;
  const int n_out = (padded && MODE != SM_FWD) ? g.n[axis] : N;
  const int nxch = (valid + TB - 1) / TB;
  const int xch = blockIdx.x % nxch;
  const int other = blockIdx.x / nxch;
  const int x0 = xch * TB;
  const size_t base = (size_t)inst * istride + (size_t)blockIdx.y * FIELDS * fstride +
                      (size_t)other * ostride + x0;

  for (int idx = tid; idx < FIELDS * N * TB; idx += nthr) {
    const int xi = idx & (TB - 1), rest = idx >> tbs, f = fN.div(rest), e = rest - f * N;
    const bool in = x0 + xi < valid && e < n_in;
    at(f, xi, e) = in ? data[base + (size_t)f * fstride + (size_t)e * estride + xi]
                      : mk((T)0, (T)0);
  }
  {
    const cplx<double>* t = gen_tw(tb, N);
    for (int i = tid; i < N; i += nthr) tw[i] = t[i];
  }
  if (MODE == SM_KIN) {  // per-axis phase tables, as strided_kernel
    const InstParams& P = tb.params[inst];
    const double dt = (dt_override > 0.0 ? dt_override : P.dt) * kin_frac;
    auto phase = [&](double k, double& re, double& im) {
      const double a = 0.5 * (k * k) * dt;
      if (itp) { re = exp(-a); im = 0.0; }
      else { double s, c; sincos_d(a, &s, &c); re = c; im = -s; }
    };
    double ore, oim;
    phase((g.ndim == 3 && axis == 2) ? tb.ktab[g.nmax + other] : 0.0, ore, oim);
    for (int i = tid; i < N + TB; i += nthr) {
      double k;
      if (i < N) k = tb.ktab[axis * g.nmax + i];
      else k = (x0 + (i - N) < g.n[0]) ? tb.ktab[x0 + (i - N)] : 0.0;
      double re, im;
      phase(k, re, im);
      if (i < N) {
        const cplx<double> v = (mk(re, im) * mk(ore, oim)) * g.inv_ntot;
        re = v.x;
        im = v.y;
      }
      tabs[2 * i] = re;
      tabs[2 * i + 1] = im;
    }
  }
  __syncthreads();

  if (MODE != SM_INV) gen_fft<T>(tile, scr, FIELDS * TB, N, LP, cap, tw, false, tid, nthr, bl);

  if (MODE == SM_KIN) {
    for (int idx = tid; idx < FIELDS * N * TB; idx += nthr) {
      const int xi2 = idx & (TB - 1), rest = idx >> tbs, f = fN.div(rest), e = rest - f * N;
      const cplx<double> pa = mk(tabs[2 * e], tabs[2 * e + 1]);
      const cplx<double> px = mk(tabs[2 * (N + xi2)], tabs[2 * (N + xi2) + 1]);
      cplx<T>& v = at(f, xi2, e);
      v = mul64(v, px * pa);
    }
    __syncthreads();
  } else if (MODE == SM_DDI) {
    const double C = tb.params[inst].cdd;
    const double s = C / ((double)g.sn[0] * (double)g.sn[1] * (double)g.sn[2]);
    for (int idx = tid; idx < N * TB; idx += nthr) {
      const int xi2 = idx & (TB - 1), e = idx >> tbs;
      const int x = x0 + xi2;
      double kx = 0.0;  // rfftfreq along x (Nyquist +), fftfreq along y and z (ddi.jl:66-68)
      if (x < valid) kx = (x == g.sn[0] / 2) ? -tb.sktab[g.sn[0] / 2] : tb.sktab[x];
      double ky, kz;
      if (axis == 2) { ky = tb.sktab[g.smax + other]; kz = tb.sktab[2 * g.smax + e]; }
      else { ky = tb.sktab[g.smax + e]; kz = 0.0; }
      cplx<T> fx = at(0, xi2, e), fy = at(1, xi2, e), fz = at(2, xi2, e);
      ddi_qmix<T>(kx, ky, kz, g.ndim, g.ddi_secular != 0, s, fx, fy, fz);
      at(0, xi2, e) = fx;
      at(1, xi2, e) = fy;
      at(2, xi2, e) = fz;
    }
    __syncthreads();
  } else if (MODE == SM_DERIV) {
    const int c = blockIdx.y;  // fft_utils.jl:4-19,25-44: component c takes i k_c / N
    for (int idx = tid; idx < N * TB; idx += nthr) {
      const int xi2 = idx & (TB - 1), e = idx >> tbs;
      const int x = x0 + xi2;
      double k = 0.0;
      if (c == 0) k = x < valid ? tb.ktab[x] : 0.0;
      else if (c == 1) k = tb.ktab[g.nmax + (axis == 1 ? e : other)];
      else if (c == 2 && g.ndim == 3) k = tb.ktab[2 * g.nmax + e];
      cplx<T>& v = at(0, xi2, e);
      v = mul64(v, mk(0.0, k * g.inv_ntot));
    }
    __syncthreads();
  } else if (MODE == SM_KE) {
    double acc = 0.0;
    for (int idx = tid; idx < FIELDS * N * TB; idx += nthr) {
      const int xi2 = idx & (TB - 1), rest = idx >> tbs, f = fN.div(rest), e = rest - f * N;
      const int x = x0 + xi2;
      if (x >= valid) continue;
      const double kx = tb.ktab[x];
      double k2 = kx * kx;
      if (axis == 2) {
        const double ky = tb.ktab[g.nmax + other], kz = tb.ktab[2 * g.nmax + e];
        k2 += ky * ky;
        k2 += kz * kz;
      } else {
        const double ky = tb.ktab[g.nmax + e];
        k2 += ky * ky;
      }
      acc += k2 * (double)abs2(at(f, xi2, e));
    }
    const double v = block_sum(acc, rsh);
    if (tid == 0) *partial_ptr(tb, RS_KIN, inst, blockIdx.x + gridDim.x * blockIdx.y) = v;
    return;
  }

  if (MODE == SM_INV || MODE == SM_KIN || MODE == SM_DDI || MODE == SM_DERIV)
    gen_fft<T>(tile, scr, FIELDS * TB, N, LP, cap, tw, true, tid, nthr, bl);

  for (int idx = tid; idx < FIELDS * N * TB; idx += nthr) {
    const int xi2 = idx & (TB - 1), rest = idx >> tbs, f = fN.div(rest), e = rest - f * N;
    if (x0 + xi2 < valid && e < n_out)
      data[base + (size_t)f * fstride + (size_t)e * estride + xi2] = at(f, xi2, e);
  }
}

// ---------------------------------------------------------------------------
// host side: tile shapes and launchers
// ---------------------------------------------------------------------------
static constexpr size_t kGenLdsSoft = 64 * 1024;   // shapes are shrunk towards this
static constexpr size_t kGenLdsHard = 160 * 1024;  // and must fit this

// Bluestein or not: the mixed-radix passes cost about N (p - 1) complex multiply-adds per
// line for each odd prime factor p > 7 (one p-term sum per output), Bluestein about
// M (1.5 log2 M + 3) (two M-point radix-4 transforms and three pointwise products) in that
// unit, but its FP64 work lines take LDS from the line tiles and its passes synchronise the
// block more often: measured on MI355X, it was slower than the mixed-radix passes at a model
// ratio of 1.4 to 1.9 (3-D sides 106, 118, 122) and 1.5 times faster at 3.8 (side 254), so it
// is taken from a ratio of 2.5 (sides 214, 218, 226, 254, and 25 sides from 314 to 508).
// Lengths whose M would exceed 4096 keep the mixed-radix passes.
int gen_blue_m(int N) {
  static const bool off = [] {
    const char* e = getenv("BEC_BLUESTEIN");
    return e && atoi(e) == 0;
  }();
  if (N < 2 || off) return 0;
  long long direct = 0;
  int rem = N;
  while (rem % 2 == 0) rem /= 2;
  for (int p = 3; p * p <= rem || rem > 1; p += 2) {
    if (p * p > rem) p = rem;
    while (rem % p == 0) {
      if (p > 7) direct += (long long)N * (p - 1);
      rem /= p;
    }
  }
  int M = 1, lg = 0;
  while (M < 2 * N - 1) {
    M *= 2;
    ++lg;
  }
  if (M > 4096) return 0;
  const long long blue = (long long)M * (3 * lg + 6) / 2;
  return 2 * direct > 5 * blue ? M : 0;
}
// scratch elements (of csz bytes) Bluestein needs for one line of length N (0: none): two
// FP64 M-point work lines after a 16-byte alignment, plus a line pitch for the rounding of
// the scratch region into whole lines
static long long blue_need(int N, int pitch, size_t csz) {
  const int M = gen_blue_m(N);
  return M ? (32LL * M + 16 + (long long)csz - 1) / (long long)csz + pitch : 0;
}

static size_t xgen_lds(int D, int NX, int AL, int rows, size_t csz, bool has_aux, long long cap) {
  return sizeof(cplx<double>) * ((size_t)NX + (AL != NX ? AL : 0) + 2 * kMaxOps * D) +
         16 * sizeof(double) +
         csz * ((size_t)D * rows * (NX + 1) + (has_aux ? (size_t)2 * rows * (AL + 1) : 0) +
                (size_t)cap);
}
// scratch elements for a full ping-pong of the largest FFT of a launch (at least one
// Bluestein line group of each length that has one)
static long long xgen_full_cap(int D, int NX, int AL, int rows, bool has_aux, size_t csz) {
  const long long t = (long long)D * rows * (NX + 1);
  const long long a = has_aux ? 2LL * rows * (AL + 1) : 0;
  return std::max({t, a, blue_need(NX, NX + 1, csz), has_aux ? blue_need(AL, AL + 1, csz) : 0LL});
}

// x sides with a built line FFT for the x pass (zero-padded DDI also needs lines of twice
// the side, and 2F+1 lines of n_x / E threads each must fit one block of 512 threads)
bool xpass_generic(const Geom& g) {
  const int nx = g.n[0];
  static constexpr int kBuilt[] = {8, 16, 24, 32, 40, 48, 56, 64, 80, 96, 112, 128, 160,
                                   192, 224, 256, 512};
  bool built = false;
  for (int b : kBuilt) built = built || b == nx;
  if (!built) return true;
  if (g.spad && fft_size_id(2 * nx) < 0) return true;
  return g.D * (nx / line_e(nx)) > 512;
}

int xgen_rows(const Geom& g) {
  // about 2048 tile elements per block; bounded for FP64 with zero-padded DDI lines (so one
  // partition serves every program and both precisions)
  const int NX = g.n[0], AL = 2 * NX;
  const int target = std::max(1, 2048 / (g.D * NX));
  int r = 1;
  while (r * 2 <= target) r *= 2;
  while (r > 1 && (g.nrows % r) != 0) r /= 2;
  while (r > 1 && xgen_lds(g.D, NX, AL, r, 16, true, xgen_full_cap(g.D, NX, AL, r, true, 16)) >
                      kGenLdsSoft)
    r /= 2;
  return r;
}

// scratch elements of an x-pass launch (0: the tile does not fit in LDS)
static long long xgen_cap(const Geom& g, int rows, size_t csz, bool has_aux) {
  const int NX = g.n[0], AL = g.sn[0];
  long long cap = xgen_full_cap(g.D, NX, AL, rows, has_aux, csz);
  const long long minc =
      std::max({(long long)(has_aux ? std::max(NX, AL) + 1 : NX + 1), blue_need(NX, NX + 1, csz),
                has_aux ? blue_need(AL, AL + 1, csz) : 0LL});
  while (cap > minc && xgen_lds(g.D, NX, AL, rows, csz, has_aux, cap) > kGenLdsHard)
    cap = std::max(minc, cap / 2);
  return xgen_lds(g.D, NX, AL, rows, csz, has_aux, cap) <= kGenLdsHard ? cap : 0;
}

struct SGenShape {
  int TB, cap;
  size_t lds;
};
static size_t sgen_lds(int N, int tb, int fields, size_t csz, int cap) {
  return sizeof(cplx<double>) * N + sizeof(double) * (2 * (N + tb + 1) + 16) +
         csz * (size_t)(fields * tb + cap) * (N + 1);
}
static SGenShape sgen_shape(int N, int fields, int csz) {
  // scratch lines (pitch N + 1) that hold one Bluestein line group
  const long long bn = blue_need(N, N + 1, (size_t)csz);
  const int bl = (int)((bn + N) / (N + 1));
  for (int tb = 16; tb >= 1; tb /= 2) {
    const int cap = std::max(fields * tb, bl);
    const size_t l = sgen_lds(N, tb, fields, csz, cap);
    if (l <= kGenLdsSoft) return {tb, cap, l};
  }
  int cap = std::max(fields, bl);
  while (cap > std::max(1, bl) && sgen_lds(N, 1, fields, csz, cap) > kGenLdsHard) --cap;
  const size_t l = sgen_lds(N, 1, fields, csz, cap);
  if (l > kGenLdsHard) return {0, 0, 0};
  return {1, cap, l};
}
int sgen_tb(int N, int fields, int csize) { return sgen_shape(N, fields, csize).TB; }

bool generic_fits(const Geom& g, int csz, int ddi) {
  if (xpass_generic(g)) {
    if (xgen_cap(g, xgen_rows(g), (size_t)csz, ddi != 0) == 0) return false;
    // the FP64 energy evaluation runs on the unpadded grid
    if (xgen_cap(g, xgen_rows(g), 16, ddi != 0) == 0) return false;
  }
  for (int a = 1; a < g.ndim; ++a) {
    if (fft_size_id(g.n[a]) < 0 && sgen_tb(g.n[a], 1, 16) == 0) return false;
    if (ddi && fft_size_id(g.sn[a]) < 0 && sgen_tb(g.sn[a], 3, 16) == 0) return false;
  }
  return true;
}

template <typename T, int D>
hipError_t xgen_launch_d(hipStream_t st, const Geom& g, const DevTables& tb, const XProgram& prog,
                         const void* src, void* dst, void* fhat, int n_active, void* phi) {
  const int rows = xpass_rows_g(g);
  const bool has_aux = (prog.stages & (XS_LOAD_PHI | XS_WRITE_FHAT)) != 0;
  const long long cap = xgen_cap(g, rows, sizeof(cplx<T>), has_aux);
  if (cap == 0 || g.nrows % rows != 0) return hipErrorInvalidConfiguration;
  if (!has_gen_tw(tb, g.n[0]) || (has_aux && !has_gen_tw(tb, g.sn[0]))) return hipErrorInvalidValue;
  const size_t lds = xgen_lds(g.D, g.n[0], g.sn[0], rows, sizeof(cplx<T>), has_aux, cap);
  if (lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void*)xgen_kernel<T, D>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const dim3 grid(g.nrows / rows, n_active);
  hipLaunchKernelGGL((xgen_kernel<T, D>), grid, dim3(kGenThreads), lds, st, g, tb, prog,
                     (const cplx<T>*)src, (cplx<T>*)dst, (cplx<T>*)fhat, (T*)phi, rows, (int)cap);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_xpass_gen(hipStream_t st, const Geom& g, const DevTables& tb,
                            const XProgram& prog, const void* src, void* dst, void* fhat,
                            int n_active, void* phi) {
#define BEC_GCALL(D) return xgen_launch_d<T, D>(st, g, tb, prog, src, dst, fhat, n_active, phi)
  switch (g.D) {
    case 1: BEC_GCALL(1);
    case 3: BEC_GCALL(3);
    case 5: BEC_GCALL(5);
    case 7: BEC_GCALL(7);
    case 9: BEC_GCALL(9);
    case 11: BEC_GCALL(11);
    case 13: BEC_GCALL(13);
    default: return hipErrorInvalidValue;
  }
#undef BEC_GCALL
}

template <typename T, int MODE, int FIELDS>
static hipError_t sgen_n(hipStream_t st, const Geom& g, const DevTables& tb, int axis, void* data,
                         int n_active, double kin_frac, double dt_override, int itp, int ddi_data,
                         int other_lines) {
  const int N = ddi_data ? g.sn[axis] : g.n[axis];
  const SGenShape s = sgen_shape(N, FIELDS, (int)sizeof(cplx<T>));
  if (s.TB == 0) return hipErrorInvalidConfiguration;
  if (!has_gen_tw(tb, N)) return hipErrorInvalidValue;
  const int valid = ddi_data ? g.nxh : g.n[0];
  const int nxch = (valid + s.TB - 1) / s.TB;
  const int ngroups = ddi_data ? 1 : g.D / FIELDS;
  const dim3 grid(nxch * other_lines, ngroups, n_active);
  if (s.lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void*)sgen_kernel<T, MODE, FIELDS>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)s.lds);
  hipLaunchKernelGGL((sgen_kernel<T, MODE, FIELDS>), grid, dim3(kGenThreads), s.lds, st, g, tb,
                     (cplx<T>*)data, axis, s.TB, s.cap, kin_frac, dt_override, itp, ddi_data);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_strided_gen(hipStream_t st, const Geom& g, const DevTables& tb, int axis,
                              int mode, void* data, int n_active, double kin_frac,
                              double dt_override, int itp, int ddi_data, int other_lines) {
#define BEC_SARGS st, g, tb, axis, data, n_active, kin_frac, dt_override, itp, ddi_data, other_lines
  if (ddi_data) {
    switch (mode) {
      case SM_FWD: return sgen_n<T, SM_FWD, 3>(BEC_SARGS);
      case SM_INV: return sgen_n<T, SM_INV, 3>(BEC_SARGS);
      case SM_DDI: return sgen_n<T, SM_DDI, 3>(BEC_SARGS);
      default: return hipErrorInvalidValue;
    }
  }
  switch (mode) {
    case SM_FWD: return sgen_n<T, SM_FWD, 1>(BEC_SARGS);
    case SM_INV: return sgen_n<T, SM_INV, 1>(BEC_SARGS);
    case SM_KIN: return sgen_n<T, SM_KIN, 1>(BEC_SARGS);
    case SM_KE: return sgen_n<T, SM_KE, 1>(BEC_SARGS);
    case SM_DERIV:
      if constexpr (sizeof(T) == 8) return sgen_n<T, SM_DERIV, 1>(BEC_SARGS);
      else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
#undef BEC_SARGS
}

template hipError_t launch_xpass_gen<float>(hipStream_t, const Geom&, const DevTables&,
                                            const XProgram&, const void*, void*, void*, int, void*);
template hipError_t launch_xpass_gen<double>(hipStream_t, const Geom&, const DevTables&,
                                             const XProgram&, const void*, void*, void*, int, void*);
template hipError_t launch_strided_gen<float>(hipStream_t, const Geom&, const DevTables&, int, int,
                                              void*, int, double, double, int, int, int);
template hipError_t launch_strided_gen<double>(hipStream_t, const Geom&, const DevTables&, int, int,
                                               void*, int, double, double, int, int, int);

}  // namespace bec
