// Types shared by the HIP kernels and the host-side workspace code.
#pragma once
#include <stdint.h>

namespace bec {

constexpr int kMaxD = 13;   // 2F+1 for F <= 6
constexpr int kMaxOps = 12;

// Per-instance physical parameters (one per independent solve in a batch).
// Device-resident; every block reads the entry of the instance it works on.
struct InstParams {
  double dt;         // time step of this instance
  double c0, c1, c_lhy;
  double c2;         // nematic (singlet-pair) coupling, 0 when the tensor step is active
  double cdd;        // dipolar strength C_dd
  double zee[kMaxD]; // Zeeman energies E_c = -p m + q m^2, m = F - c
  double zee_shift;  // min_c E_c (subtracted in imaginary time)
  double gS[7];      // channel couplings g_S, S = 0, 2, ..., 2F (general-F tensor step)
  double loss[kMaxD]; // loss rate gamma_c + L3 per component (real time only)
  double raman[5];   // RamanCoupling: Omega_R, delta, k_eff per axis (raman.jl, types.jl:179-183)
  int active;
  int has_loss;      // 0: no loss step (LossParams absent, or gamma_dr and L3 < 1e-30)
  int has_raman;     // 1: the Raman step runs (RamanCoupling present)
};

enum OpKind : int {
  OP_DIAG = 1,      // _diagonal_step_svec! (dt*frac)
  OP_SPINMIX = 2,   // apply_spin_mixing_step! (dt*frac)
  OP_DDI_ROT = 3,   // Euler rotation by the DDI potential (dt*frac)
  OP_SCALE = 4,     // multiply by 1/sqrt(norm2) of the instance (ITP renormalisation)
  OP_TENSOR = 5,    // apply_tensor_interaction_step! (dt*frac)
  OP_NEMATIC = 6,   // apply_nematic_step! (dt*frac)
  OP_LOSS = 7,      // apply_loss_step! (dt*frac), real time
  OP_RAMAN = 8,     // apply_raman_step! (dt*frac)
};

struct Op {
  int kind;
  int pad_;
  double frac;
};

// Program executed by the x-axis pass kernel on a tile of whole x-rows holding
// all 2F+1 components (so every per-site operation sees the full spinor).
enum XStage : int {
  XS_INV_FFT = 1 << 0,     // inverse x FFT on load (end of a kinetic step)
  XS_LOAD_PHI = 1 << 1,    // read DDI Phi_hat half rows, C2R along x
  XS_OPS_A = 1 << 2,       // per-site op list A
  XS_KIN_1D = 1 << 3,      // 1-D grid: forward FFT, kinetic phase, inverse FFT in LDS
  XS_OPS_B = 1 << 4,       // per-site op list B (after the 1-D kinetic step)
  XS_WRITE_FHAT = 1 << 5,  // spin density -> R2C along x -> F_hat half rows
  XS_NORM = 1 << 6,        // per-block partial of sum |psi|^2 (after ops)
  XS_FWD_FFT = 1 << 7,     // forward x FFT before the store
  XS_NO_STORE = 1 << 8,    // do not write psi back
  XS_KE_1D = 1 << 10,      // 1-D kinetic energy partial sum k^2 |psi_hat|^2 after forward FFT
  XS_STORE_PHI = 1 << 12,  // after XS_LOAD_PHI: write Phi(x) (3 real fields) for the site kernel
  XS_NO_LOAD = 1 << 13,    // psi is not read (with XS_NO_STORE: a pass over the DDI spectra only)
};

struct XProgram {
  int stages;
  int nA, nB;
  int itp;            // imaginary time
  double dt_override; // > 0: use instead of the per-instance dt
  double kin_frac;    // 1-D kinetic step dt fraction
  int deriv;          // KIN_1D multiplies by i k / N (spectral derivative) instead of the phase
  int nrep;           // > 1: the ops A / KIN_1D / ops B / NORM stages run nrep times on the tile
                      // (one block per instance holds the whole 1-D line); with XS_NORM every
                      // repetition after the first starts with the renormalisation by the
                      // previous one's norm
  Op A[kMaxOps];
  Op B[kMaxOps];
};

// Program of the lane-per-site kernel (heavy per-site operators of large spins).
struct SProgram {
  int n;
  int itp;
  double dt_override;
  Op ops[kMaxOps];
};

// Reduction slots written per block into the partials buffer.
enum RedSlot : int {
  RS_NORM = 0,
  RS_KIN = 1,
  RS_TRAP = 2,
  RS_ZEE = 3,
  RS_C0 = 4,
  RS_C1 = 5,
  RS_LHY = 6,
  RS_DDI = 7,
  RS_MZ = 8,
  RS_PHIMAX = 9,
  RS_TENSOR = 10,
  RS_RAMAN = 11,
  RS_COUNT = 12,
};

}  // namespace bec

// Enumeration of the Clebsch-Gordan table used by the tensor step: for every
// M = -2F..2F, every unordered pair m1 <= m2 with m1 + m2 = M, every even
// S >= |M| up to 2F: <F m1; F m2 | S M>.  Shared by host (table build) and
// device (compile-time offsets).
namespace tens {
__host__ __device__ constexpr int m1lo(int F, int M) { return M - F > -F ? M - F : -F; }
__host__ __device__ constexpr int m1hi(int M) { return M >= 0 ? M / 2 : -((-M + 1) / 2); }
__host__ __device__ constexpr int npairs(int F, int M) { return m1hi(M) - m1lo(F, M) + 1; }
__host__ __device__ constexpr int s0(int M) { return ((M < 0 ? -M : M) + 1) / 2 * 2; }
__host__ __device__ constexpr int ns(int F, int M) { return (2 * F - s0(M)) / 2 + 1; }
__host__ __device__ constexpr int base(int F, int M) {
  int b = 0;
  for (int q = -2 * F; q < M; ++q) b += npairs(F, q) * ns(F, q);
  return b;
}
__host__ __device__ constexpr int size(int F) { return base(F, 2 * F + 1); }
// Channel-folded kernel K_M[i][j] = sum_S CG(i | S M) g_S CG(j | S M) over the unordered
// pairs i, j of each M, stored M by M as full npairs x npairs blocks (kfold_off(F, M) is the
// start of block M).
__host__ __device__ constexpr int kfold_off(int F, int M) {
  int b = 0;
  for (int q = -2 * F; q < M; ++q) b += npairs(F, q) * npairs(F, q);
  return b;
}
__host__ __device__ constexpr int kfold_size(int F) { return kfold_off(F, 2 * F + 1); }
constexpr int kKfoldMax = 413;  // kfold_size(6)
__host__ __device__ constexpr int off(int F, int M, int m1, int S) {
  return base(F, M) + (m1 - m1lo(F, M)) * ns(F, M) + (S - s0(M)) / 2;
}

// Compile-time Clebsch-Gordan coefficients (Racah's closed form, Condon-Shortley phase),
// so the device tensor step takes them as instruction literals instead of loads.
__host__ __device__ constexpr double cfact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= (double)i;
  return r;
}
__host__ __device__ constexpr double csqrt(double x) {  // Newton from above, monotone
  if (!(x > 0.0)) return 0.0;
  double r = x < 1.0 ? 1.0 : x;
  for (int i = 0; i < 4000; ++i) {
    const double n = 0.5 * (r + x / r);
    if (!(n < r)) break;
    r = n;
  }
  return r;
}
__host__ __device__ constexpr double cg(int j1, int m1, int j2, int m2, int J, int M) {
  if (m1 + m2 != M || m1 > j1 || -m1 > j1 || m2 > j2 || -m2 > j2 || M > J || -M > J) return 0.0;
  if (J > j1 + j2 || J < (j1 > j2 ? j1 - j2 : j2 - j1)) return 0.0;
  const double pre = (2.0 * J + 1.0) * cfact(J + j1 - j2) * cfact(J - j1 + j2) *
                     cfact(j1 + j2 - J) / cfact(j1 + j2 + J + 1) * cfact(J + M) * cfact(J - M) *
                     cfact(j1 - m1) * cfact(j1 + m1) * cfact(j2 - m2) * cfact(j2 + m2);
  int klo = 0;
  if (j2 - J - m1 > klo) klo = j2 - J - m1;
  if (j1 - J + m2 > klo) klo = j1 - J + m2;
  int khi = j1 + j2 - J;
  if (j1 - m1 < khi) khi = j1 - m1;
  if (j2 + m2 < khi) khi = j2 + m2;
  double sum = 0.0;
  for (int k = klo; k <= khi; ++k) {
    const double d = cfact(k) * cfact(j1 + j2 - J - k) * cfact(j1 - m1 - k) * cfact(j2 + m2 - k) *
                     cfact(J - j2 + m1 + k) * cfact(J - j1 - m2 + k);
    sum += (k % 2 ? -1.0 : 1.0) / d;
  }
  return csqrt(pre) * sum;
}
// Wigner small-d at beta = pi/2: d^j_{m'm}(pi/2) = <j m'| exp(-i pi/2 F_y) |j m>
// = 2^-j sqrt((j+m')!(j-m')!(j+m)!(j-m)!) sum_k (-1)^(k-m+m') / ((j+m-k)! k! (j-k-m')! (k-m+m')!).
__host__ __device__ constexpr double wigner_d_half_pi(int j, int mp, int m) {
  double s = 0.0;
  for (int k = 0; k <= 2 * j; ++k) {
    if (j + m - k < 0 || j - k - mp < 0 || k - m + mp < 0) continue;
    const double den = cfact(j + m - k) * cfact(k) * cfact(j - k - mp) * cfact(k - m + mp);
    s += ((k - m + mp) % 2 != 0 ? -1.0 : 1.0) / den;
  }
  double p = 1.0;
  for (int i = 0; i < j; ++i) p *= 0.5;
  return p * csqrt(cfact(j + mp) * cfact(j - mp) * cfact(j + m) * cfact(j - m)) * s;
}
// d^F(pi/2) in component order (c <-> m = F - c), row-major [c'][c].  With
// U = exp(-i pi/2 F_z) d^F(pi/2), F_y = U F_z U^dagger: U holds eigenvectors of F_y.
template <int F>
struct DHalf {
  double v[(2 * F + 1) * (2 * F + 1)];
  __host__ __device__ constexpr DHalf() : v() {
    for (int cp = 0; cp <= 2 * F; ++cp)
      for (int c = 0; c <= 2 * F; ++c) {
        const double x = wigner_d_half_pi(F, F - cp, F - c);
        v[cp * (2 * F + 1) + c] = (x < 1e-15 && x > -1e-15) ? 0.0 : x;
      }
  }
};

// Rows of d = d^F(pi/2) and of d^T are even or odd under m -> -m (column c -> D-1-c):
// d_{m',-m}(pi/2) = (-1)^(F+m') d_{m',m}(pi/2), and d^T = P d P with P = diag((-1)^m).  The
// parity of each row is read off the table (+1 even, -1 odd, 0 neither, which keeps the
// row's full sum; equal within 1e-14, the closed form's rounding).
template <int F, bool TRANS>
struct DHalfParity {
  int s[2 * F + 1];
  __host__ __device__ constexpr DHalfParity() : s() {
    constexpr int D = 2 * F + 1;
    const DHalf<F> d{};
    for (int i = 0; i < D; ++i) {
      bool ev = true, od = true;
      for (int j = 0; j < D; ++j) {
        const double a = TRANS ? d.v[j * D + i] : d.v[i * D + j];
        const double b = TRANS ? d.v[(D - 1 - j) * D + i] : d.v[i * D + D - 1 - j];
        if (a - b > 1e-14 || b - a > 1e-14) ev = false;
        if (a + b > 1e-14 || -(a + b) > 1e-14) od = false;
      }
      s[i] = ev ? 1 : (od ? -1 : 0);
    }
  }
};

// The table in off() order, with |c| < 1e-15 set to 0 (clebsch_gordan.jl:174-175).
template <int F>
struct CGTab {
  double v[size(F)];
  __host__ __device__ constexpr CGTab() : v() {
    for (int M = -2 * F; M <= 2 * F; ++M)
      for (int m1 = m1lo(F, M); m1 <= m1hi(M); ++m1)
        for (int S = s0(M); S <= 2 * F; S += 2) {
          const double c = cg(F, m1, F, M - m1, S, M);
          v[off(F, M, m1, S)] = (c < 1e-15 && c > -1e-15) ? 0.0 : c;
        }
  }
};
}  // namespace tens
