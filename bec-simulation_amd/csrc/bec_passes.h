// Device building blocks shared by the x pass and the fused x/site pass: block
// reductions into per-block partials, the trap potential at a site, and the auxiliary
// DDI lines (C2R input of the potential spectra, R2C output of the spin density) that
// both kernels keep beside their LDS tile of x rows.
#pragma once
#include <hip/hip_runtime.h>

#include "bec_device.h"
#include "bec_kernels.h"

namespace bec {

// Threads per x-pass block are capped so that the per-site operators of large spins
// (13-component Euler rotations and the tensor step) keep their spinors in VGPRs.
template <int D>
constexpr int xpass_max_threads() { return 512; }
// Values per FFT thread along x (the line's radix plan: 8, 12, 20 or 28).  With E <= D
// every tile has at most one site per thread.
template <int NX, int D>
constexpr int xpass_e() { return fft_e<NX>(); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Deterministic block sum (result on thread 0).  sh: >= 16 doubles of LDS.
__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += sh[i];
  }
  return r;
}
__device__ inline double block_max(double v, double* sh) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r = fmax(r, sh[i]);
  }
  return r;
}

__device__ __forceinline__ double* partial_ptr(const DevTables& tb, int slot, int inst, int blk) {
  return tb.partials + ((size_t)slot * tb.maxB + inst) * tb.max_blocks + blk;
}

// Twiddle table exp(-2 pi i m / N), m < N, of a built line length.
template <int N>
__device__ __forceinline__ const cplx<double>* tw_table(const DevTables& tb) {
  static_assert(fft_size_id(N) >= 0, "no FFT of this length is built");
  return reinterpret_cast<const cplx<double>*>(tb.tw[fft_size_id(N)]);
}

// a / n for a grid side n: a shift when n is a power of two (POW2: known to be, as in the
// fused x/site kernel, which only runs such grids), else a division (n is uniform over
// the launch, so the branch is too)
template <bool POW2 = false>
__device__ __forceinline__ int div_side(int a, int n) {
  if constexpr (POW2) return a >> (__ffs(n) - 1);
  else return (n & (n - 1)) == 0 ? a >> (__ffs(n) - 1) : a / n;
}

// Trap potential at a flat site index (potentials.jl:5-15 via per-axis omega^2 x^2 tables,
// or an explicit potential array).
template <bool POW2 = false>
__device__ __forceinline__ double trap_site(const Geom& g, const DevTables& tb, int inst,
                                            long long site) {
  if (tb.Varr) return tb.Varr[(size_t)inst * g.nsites + site];
  // an instance has fewer than 2^31 sites
  const int sidx = (int)site;
  const int row = div_side<POW2>(sidx, g.n[0]);
  const int x = sidx - row * g.n[0];
  const int z = div_side<POW2>(row, g.n[1]), y = row - z * g.n[1];
  const double* ax = tb.axtab + (size_t)inst * 3 * g.nmax;
  double v = ax[x];
  if (g.ndim >= 2) v += ax[g.nmax + y];
  if (g.ndim >= 3) v += ax[2 * g.nmax + z];
  return 0.5 * v;
}

// raman.jl:19-23: H_R = delta Fz + Omega_R/2 (e^{i k.r} F+ + e^{-i k.r} F-) at a flat site
// index is phi . F with phi = (Omega_R cos k.r, -Omega_R sin k.r, delta)
// (F+- = Fx +- i Fy), so exp(-i H_R dt) is the Euler rotation by phi; k.r is summed over the
// axes in order in FP64.
template <bool POW2 = false>
__device__ __forceinline__ void raman_phi(const Geom& g, const DevTables& tb, const InstParams& P,
                                          long long site, double& phx, double& phy, double& phz) {
  const int sidx = (int)site;
  const int row = div_side<POW2>(sidx, g.n[0]);
  const int x = sidx - row * g.n[0];
  const int z = div_side<POW2>(row, g.n[1]), y = row - z * g.n[1];
  double kr = P.raman[2] * tb.xtab[x];
  if (g.ndim >= 2) kr += P.raman[3] * tb.xtab[g.nmax + y];
  if (g.ndim >= 3) kr += P.raman[4] * tb.xtab[2 * g.nmax + z];
  double s, c;
  sincos_d(kr, &s, &c);
  phx = P.raman[0] * c;
  phy = -P.raman[0] * s;
  phz = P.raman[1];
}

// LDS tiles are padded with one slot per 8 elements: element e lives at p8(e) = e + e/8,
// so the stride-8 writes of the first Stockham pass do not pile onto one bank.  Lines
// start at multiples of their length (a multiple of 8), so p8(l*N + i) = p8(l*N) + p8(i).
__device__ __forceinline__ int p8(int e) { return e + (e >> 3); }

// Row of the DDI spectra (y, z) -> y + z*sn[1] for grid row `row` = y + z*n[1].
template <bool POW2 = false>
__device__ __forceinline__ int spec_row(const Geom& g, int row) {
  const int z = div_side<POW2>(row, g.n[1]);
  return (row - z * g.n[1]) + z * g.sn[1];
}

// C2R input along x of the three DDI potential spectra for the ROWS tile rows from row0
// (reference irfft semantics, ddi.jl:108-136: bins 0..AL/2 with the imaginary parts of DC
// and Nyquist dropped, Hermitian extension).  Two-for-one: auxiliary line 2r holds
// Phi_x + i Phi_y, line 2r+1 holds Phi_z (its inverse FFT is real).
// Element e = r*AL + x of the lines in two halves: ddi_phi_fetch reads its three spectral
// bins (so a caller can issue the loads early), ddi_phi_place writes the line entries.
template <typename T, bool POW2 = false>
__device__ __forceinline__ void ddi_phi_fetch(const Geom& g, const cplx<T>* __restrict__ fhat,
                                              size_t fb, int row0, int AL, int e, cplx<T>* v) {
  const int r = e / AL, x = e - r * AL;
  const size_t rowoff = (size_t)spec_row<POW2>(g, row0 + r) * g.nxh_pitch;
  const int kk = x <= AL / 2 ? x : AL - x;
  v[0] = fhat[fb + rowoff + kk];
  v[1] = fhat[fb + g.fhat_field_stride + rowoff + kk];
  v[2] = fhat[fb + 2 * g.fhat_field_stride + rowoff + kk];
}
template <typename T>
__device__ __forceinline__ void ddi_phi_place(int AL, int e, const cplx<T>* v, cplx<T>* aux) {
  const int r = e / AL, x = e - r * AL;
  const int kk = x <= AL / 2 ? x : AL - x;
  cplx<T> yx = v[0], yy = v[1], yz = v[2];
  if (kk == 0 || kk == AL / 2) {
    yx.y = 0; yy.y = 0; yz.y = 0;
  } else if (x > AL / 2) {
    yx = conjc(yx); yy = conjc(yy); yz = conjc(yz);
  }
  aux[p8((2 * r) * AL + x)] = mk(yx.x - yy.y, yx.y + yy.x);
  aux[p8((2 * r + 1) * AL + x)] = yz;
}
template <typename T>
__device__ __forceinline__ void ddi_lines_in(const Geom& g, const cplx<T>* __restrict__ fhat,
                                             size_t fb, int row0, int ROWS, int AL, cplx<T>* aux,
                                             int tid, int nthr) {
  for (int e = tid; e < ROWS * AL; e += nthr) {
    cplx<T> v[3];
    ddi_phi_fetch<T>(g, fhat, fb, row0, AL, e, v);
    ddi_phi_place<T>(AL, e, v, aux);
  }
}

// FFT of the 2*ROWS auxiliary lines of length AL (NX, or 2 NX with zero padding), in
// rounds of nthr / (AL/E) lines (E values per thread).  Every thread of the block calls it.
template <typename T, int NX, bool INV>
__device__ __forceinline__ void ddi_lines_fft(const DevTables& tb, cplx<T>* aux, int ROWS, int AL,
                                              int tid, int nthr, const cplx<double>* tw) {
  if (AL == NX) {
    constexpr int E = fft_e<NX>();
    constexpr int TEAM = NX / E;
    const int lpr = nthr / TEAM, l = tid / TEAM, t = tid - l * TEAM;
    for (int l0 = 0; l0 < 2 * ROWS; l0 += lpr) {
      const int ln = l0 + l;
      const bool act = l < lpr && ln < 2 * ROWS;
      fft_team<T, NX, INV, E, true>(aux + p8((act ? ln : 0) * NX), 1, t, act, tw);
    }
  } else {
    if constexpr (fft_size_id(2 * NX) >= 0) {  // padded lines of a built length only
      constexpr int E2 = fft_e<2 * NX>();
      constexpr int TEAM2 = 2 * NX / E2;
      const cplx<double>* tw2 = tw_table<2 * NX>(tb);
      const int lpr = nthr / TEAM2, l2 = tid / TEAM2, t2 = tid - l2 * TEAM2;
      for (int l0 = 0; l0 < 2 * ROWS; l0 += lpr) {
        const int ln = l0 + l2;
        const bool act = l2 < lpr && ln < 2 * ROWS;
        fft_team<T, 2 * NX, INV, E2, true>(aux + p8((act ? ln : 0) * 2 * NX), 1, t2, act, tw2);
      }
    }
  }
}

// Zero padding of the spin-density lines: entries NX..2NX-1 of every auxiliary line.
template <typename T, int NX>
__device__ __forceinline__ void ddi_lines_zero_upper(cplx<T>* aux, int ROWS, int AL, int tid,
                                                     int nthr) {
  for (int e = tid; e < 2 * ROWS * NX; e += nthr) {
    const int l = e / NX, x = NX + (e - l * NX);
    aux[p8(l * AL + x)] = mk((T)0, (T)0);
  }
}

// R2C output along x of the spin density (after the forward FFT of the auxiliary lines):
// the two-for-one split of line 2r into the spectra of F_x and F_y, line 2r+1 is F_z's;
// half rows of AL/2 + 1 bins into the DDI spectra buffer.
template <typename T, bool POW2 = false>
__device__ __forceinline__ void ddi_lines_out(const Geom& g, cplx<T>* __restrict__ fhat, size_t fb,
                                              int row0, int ROWS, int AL, const cplx<T>* aux,
                                              int tid, int nthr) {
  const int NH = AL / 2 + 1;
  for (int e = tid; e < ROWS * NH; e += nthr) {
    const int r = e / NH, k = e - r * NH;
    const cplx<T> z1 = aux[p8((2 * r) * AL + k)];
    const cplx<T> z1n = conjc(aux[p8((2 * r) * AL + (k == 0 ? 0 : AL - k))]);
    const cplx<T> sum = z1 + z1n, dif = z1 - z1n;
    const size_t o = fb + (size_t)spec_row<POW2>(g, row0 + r) * g.nxh_pitch + k;
    fhat[o] = mk(sum.x * (T)0.5, sum.y * (T)0.5);
    fhat[o + g.fhat_field_stride] = mk(dif.y * (T)0.5, -dif.x * (T)0.5);
    fhat[o + 2 * g.fhat_field_stride] = aux[p8((2 * r + 1) * AL + k)];
  }
}

// ddi.jl:7-82 and :108-136 at one spectral point: Q_ab = k_a k_b / k^2 - delta_ab / 3 with
// Q(0) = 0 (secular: Q_zz = k_z^2 / k^2 - 1/3, Q_xx = Q_yy = -Q_zz / 2, no off-diagonal
// terms), then (F_x, F_y, F_z) <- s Q (F_x, F_y, F_z) in FP64 with one rounding per output
// (s = c_dd / number of spectral points, the inverse transform's normalisation).  Axes
// above the grid's dimension contribute k = 0.
template <typename T>
__device__ __forceinline__ void ddi_qmix(double kx, double ky, double kz, int ndim, bool secular,
                                         double s, cplx<T>& fx, cplx<T>& fy, cplx<T>& fz) {
  double k2 = kx * kx;
  if (ndim >= 2) k2 += ky * ky;
  if (ndim >= 3) k2 += kz * kz;
  double qxx = 0, qyy = 0, qzz = 0, qxy = 0, qxz = 0, qyz = 0;
  if (k2 != 0.0) {
    const double ik = 1.0 / k2;
    if (secular) {
      const double q = kz * kz * ik - 1.0 / 3.0;
      qzz = q; qxx = -q / 2.0; qyy = -q / 2.0;
    } else {
      qxx = kx * kx * ik - 1.0 / 3.0;
      qyy = ky * ky * ik - 1.0 / 3.0;
      qzz = kz * kz * ik - 1.0 / 3.0;
      qxy = kx * ky * ik;
      qxz = kx * kz * ik;
      qyz = ky * kz * ik;
    }
  }
  const cplx<T> ax = fx, ay = fy, az = fz;
  auto mix = [&](double a, double b, double c) -> cplx<T> {
    const double re = a * (double)ax.x + b * (double)ay.x + c * (double)az.x;
    const double im = a * (double)ax.y + b * (double)ay.y + c * (double)az.y;
    return mk((T)(re * s), (T)(im * s));
  };
  fx = mix(qxx, qxy, qxz);
  fy = mix(qxy, qyy, qyz);
  fz = mix(qxz, qyz, qzz);
}

}  // namespace bec
